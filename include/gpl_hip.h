/*
 * gpl_hip.h -- C ABI of the MI355X ePlace Nesterov engine (libgpl_hip.so).
 *
 * Plain C, no C++ or torch types across the boundary, so a Rust host binds it with
 * `extern "C"` unchanged.  Every entry point returns an int32 status:
 *   0                 GPL_OK
 *   negative          API (-1 argument, -3 state, -4 refused) or HIP (-2) error
 *   positive          the reference's GPL message id the run stopped with
 *                     (304, 305, 306, 307: the divergence codes of nesterov.rs:2525, 2788,
 *                      2798, 2707, 3048)
 * The text of the last error of a handle is read with gpl_last_error().
 *
 * Ownership: device memory belongs to the context; host pointers passed in are borrowed for the
 * duration of the call only.  One context per host thread; each context owns one HIP stream.
 *
 * Reference interfaces each entry point replaces (paths relative to the reference checkout):
 *   gpl_design_*            main.rs:44-214 (the ODB reader) -> a seeded synthetic generator,
 *                           placer_base.rs:301-648, initial_place.rs:420-446
 *   gpl_init_nesterov_place nesterov.rs:787  init_nesterov_place
 *   gpl_init_nesterov_loop  nesterov.rs:4123 init_nesterov_loop (reference HPWL + np_init)
 *   gpl_run_nesterov_loop   nesterov.rs:4131 run_nesterov_loop
 *   gpl_do_nesterov_place   nesterov.rs:4115 do_nesterov_place
 *   gpl_continue_nesterov_place nesterov.rs:4243 continue_nesterov_place
 *   gpl_placed_locations    nesterov.rs:4099 placed_locations
 *   kernel-level entries    the per-function loops of nesterov.rs:2139-2507, 2718-2772
 */
#ifndef GPL_HIP_H
#define GPL_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GPL_OK 0
#define GPL_ERR_ARG (-1)
#define GPL_ERR_HIP (-2)
#define GPL_ERR_STATE (-3)
#define GPL_ERR_REFUSED (-4)

/* Cell flags (gpl_problem.cell_flags). */
#define GPL_CELL_INSTANCE 1u /* GCell::is_instance (nesterov.rs:161); a filler otherwise */
#define GPL_CELL_MACRO 2u    /* GCell::macro_ (nesterov.rs:116) */
#define GPL_CELL_LOCKED 4u   /* GCell::locked (nesterov.rs:127) */

/* ------------------------------------------------------------------------------------------ */
/* Options                                                                                    */
/* ------------------------------------------------------------------------------------------ */

/* `Opts` (nesterov.rs:35-82), `IpVars` (initial_place.rs:19-32) and the construction `Vars`
 * (placer_base.rs:209-216), with the reference's defaults from gpl_opts_default(). */
typedef struct gpl_opts {
    float density;                 /* -density, 0.7 */
    int32_t uniform_density;       /* -density uniform */
    int32_t bin_cnt;               /* -bin_grid_count; 0: derived per region (nesterov.rs:1847) */
    float init_density_penalty;    /* 8e-5 */
    float init_wirelength_coef;    /* 0.25 */
    float min_phi_coef;            /* 0.95 */
    float max_phi_coef;            /* 1.05 */
    float overflow;                /* target overflow, 0.1 */
    float reference_hpwl;          /* 0: derived (nesterov.rs:96) */
    int32_t max_iter;              /* 5000 */
    int32_t seed;                  /* -random_seed, 1 */
    float perturb_dist;            /* -perturb_dist, -1: min(0.5 um, row) */
    int32_t disable_revert_if_diverge;
    int32_t pad_left;
    int32_t pad_right;
    int32_t skip_io;
    int32_t disable_pin_density_adjust;
    int32_t initial_place_max_iter;   /* 20 */
    int32_t initial_place_max_fanout; /* 200 */
    int32_t initial_place_min_diff_length; /* 1500 */
    int32_t initial_place_max_solver_iter; /* 100 */
    float initial_place_net_weight_scale;  /* 800 */
    int32_t skip_initial_place;
    /* Routability-driven placement (-routability_driven: Opts.routability, routability_end_overflow
     * and routability_snapshot_overflow, nesterov.rs:50-54, 77-78; RouteVars, route_base.rs:12-46). */
    int32_t routability;                      /* 0: off */
    float routability_end_overflow;           /* -routability_check_overflow, 0.3 */
    float routability_snapshot_overflow;      /* -routability_snapshot_overflow, 0.6 */
    float routability_target_rc;              /* -routability_target_rc_metric, 1.01 */
    float routability_inflation_ratio_coef;   /* 2 */
    float routability_max_inflation_ratio;    /* 3 */
    float routability_max_density;            /* 0.99 */
    float routability_min_congestion;         /* -routability_min_congestion_for_inflation, 0.95 */
    float routability_max_inflation_total;    /* 0.30 of the instances' area */
    float routability_net_weight_max;         /* 1: no net reweighting */
    float routability_congested_nets_percentage; /* 1 */
    float routability_rc_k[4];                /* -routability_rc_coefficients, 1 1 0 0 */
    int32_t routability_max_passes;           /* the pass count that ends routability (GPL-0091), 50 */
    /* The RUDY grid the router supplies (RudyInputs, route_base.rs:48-55), for the synthetic
     * designs: tiles of (die width / route_gcells) DBU from the die's lower left, route_tracks
     * horizontal and vertical tracks per tile (two routing layers of pitch tile / route_tracks). */
    int32_t route_gcells;                     /* tiles along x, 64 */
    int32_t route_tracks;                     /* 8 */
} gpl_opts;

void gpl_opts_default(gpl_opts* o);

/* ------------------------------------------------------------------------------------------ */
/* Synthetic design (replaces the ODB reader; no database is available)                      */
/* ------------------------------------------------------------------------------------------ */

/* A seeded netlist generator: movable standard cells one row tall, fixed IO ports on the die
 * perimeter, nets of geometric degree, optional high-fanout nets, fixed macros and region
 * groups.  Every field is an input; 0 picks the documented default. */
typedef struct gpl_gen_config {
    int32_t num_cells;        /* movable standard cells (outside region groups) */
    int32_t num_nets;         /* nets of degree 2 + Geom(degree_p), capped */
    double degree_p;          /* geometric success probability */
    int32_t degree_cap;
    int32_t num_hf_nets;      /* extra high-fanout nets */
    int32_t hf_min_pins;
    int32_t hf_max_pins;
    int32_t num_io;           /* fixed ports, evenly on the die perimeter */
    double utilization;       /* movable area / free core area */
    double macro_area_frac;   /* fraction of the core covered by fixed macros */
    int32_t macro_rows_min;
    int32_t macro_rows_max;
    int32_t row_height;       /* DBU, 2000 */
    int32_t site_width;       /* DBU, 200 */
    int32_t width_min_sites;  /* 1 */
    int32_t width_max_sites;  /* 8 */
    int32_t core_rows;        /* 0: square core from the utilization */
    int32_t core_sites;
    int32_t dbu;              /* DBU per micron, 1000 */
    int32_t locality_window;  /* a net's cells come from a window of this many cell indices */
    uint32_t seed;
    /* Region groups (all zero for none): group g has group_cells[g] movable cells in its own
     * rectangular region; regions are laid out left to right along the bottom of the core. */
    int32_t num_groups;
    const int32_t* group_cells;
} gpl_gen_config;

typedef struct gpl_design gpl_design;

/* Builds the design records, the placer (placer_base_common + one placer_base per region,
 * placer_base.rs:301, 430) and the placer pins (initial_place.rs:100). */
int32_t gpl_design_generate(const gpl_gen_config* cfg, const gpl_opts* o, gpl_design** out);
void gpl_design_destroy(gpl_design* d);

/* B2B + BiCGSTAB initial placement (initial_place.rs:420 do_bicgstab_place), host side. */
int32_t gpl_design_initial_place(gpl_design* d);
/* The same rounds with the pin flags, the B2B system (createSparseMatrix, initial_place.rs:264-355;
 * its CSR, bicgstab.rs:29-69) and the BiCGSTAB solves (bicgstab.rs:72-197) on the GPU, in the host's
 * evaluation orders: the placement it leaves is the host's bit for bit. */
int32_t gpl_design_initial_place_gpu(gpl_design* d);

/* Design statistics for reports: counts and areas. */
typedef struct gpl_design_info {
    int32_t dbu;
    int32_t die[4];
    int32_t core[4];
    int32_t num_insts;        /* placer instances (movable + fixed) */
    int32_t num_movable;
    int32_t num_fixed;
    int32_t num_nets;
    int32_t num_pins;
    int32_t num_regions;
    int64_t movable_area;
    int64_t fixed_area;
    int64_t initial_hpwl;     /* initial_place.rs:381 hpwl of the database boxes */
    int32_t ip_rounds;        /* initial-placement rounds run */
    float ip_error;           /* the last round's larger residual */
} gpl_design_info;
int32_t gpl_design_get_info(const gpl_design* d, gpl_design_info* out);

/* Host helpers the construction uses, exported so their rules can be pinned by the reference's
 * known-answer tests: mt19937 outputs (normal.rs:38-120), boost uniform_int draws
 * (normal.rs:127; n draws from one engine, draw i in [lo[i], hi[i]]), boost normal draws
 * (normal.rs:120-214: n draws of normal(mean, sigma) from mt19937(seed), each with a fresh
 * distribution as the perturbation draws them; returns the engine's next word after them, so
 * the count of words the draws took is checked too), the x87 wavenumbers (nesterov.rs:351, 359), the fixed-area union
 * (nesterov.rs:1901) and roundDownToPowerOfTwo (nesterov.rs:771).  `rects` holds 4 int32 per
 * rectangle (lx, ly, ux, uy). */
void gpl_util_mt19937(uint32_t seed, int32_t n, uint32_t* out);
void gpl_util_uniform_int(uint32_t seed, int32_t n, const int32_t* lo, const int32_t* hi, int32_t* out);
uint32_t gpl_util_normal(uint32_t seed, double mean, double sigma, int32_t n, double* out);
float gpl_util_wavenumber_x(int32_t i, int32_t nx);
float gpl_util_wavenumber_y(int32_t i, int32_t ny, float bin_x, float bin_y);
int64_t gpl_util_union_area(const int32_t* rects, int32_t n);
uint32_t gpl_util_round_down_pow2(uint32_t x);

/* The byte sizes of the ABI's structs as this library was compiled (no reference counterpart: a
 * binding's layout check), in the order gpl_opts, gpl_gen_config, gpl_design_info, gpl_region,
 * gpl_problem, gpl_stats, gpl_iter_row, gpl_kernel_time; returns how many there are. */
int32_t gpl_util_struct_sizes(int64_t* out, int32_t cap);

/* main.rs:24-41 (updateDbGCells, the job's report of the design HPWL): each placed instance's
 * lower-left (the triples gpl_placed_locations returns: placer instance, lx, ly) written into the
 * design's instance boxes, marked placed; *hpwl_out the HPWL of the nets' terminal boxes afterwards
 * (db_router.rs:98-112's GPL-1018 measure: an instance terminal at its instance's centre plus its
 * offset, a port its box). */
int32_t gpl_design_write_locations(gpl_design* d, const int32_t* loc, int32_t n, int64_t* hpwl_out);
/* The job report's `bases` (main.rs:362-364): per placer base 4 int32 -- group (-1 the top level),
 * instances placed, fixed, dummies; returns the number of bases. */
int32_t gpl_design_bases(const gpl_design* d, int32_t* out, int32_t cap);
/* The job's write_def (main.rs:357-358, libodb's DEF writer there): the design as it stands as a
 * DEF 5.8 file -- die, rows, components (PLACED / FIXED), IO pins, nets.  The synthetic designs have
 * no library, so masters are named by size and terminals by their order on the instance. */
int32_t gpl_design_write_def(const gpl_design* d, const char* path);

/* ------------------------------------------------------------------------------------------ */
/* The Nesterov placer problem: the state init_nesterov_place (nesterov.rs:787) leaves        */
/* ------------------------------------------------------------------------------------------ */

/* One region (`Nb`, nesterov.rs:451): its bin grid, its cells (a contiguous range of the
 * problem's cells: its instances in placer order, then its fillers) and its fixed area. */
typedef struct gpl_region {
    int32_t lx, ly, ux, uy;          /* bin grid box (BinGrid, nesterov.rs:260) */
    int32_t cnt_x, cnt_y;
    double size_x, size_y;
    float target_density;
    float uniform_target_density;
    int32_t cell_begin, cell_end;    /* [begin, end) into the problem's cells */
    int32_t num_fillers;             /* the last num_fillers cells of the range */
    int32_t filler_dx, filler_dy;
    int64_t std_insts_area;
    int64_t macro_insts_area;
    int64_t non_place_insts_area;
    int64_t white_space_area;
    int64_t movable_area;
    int64_t total_filler_area;
    const int64_t* non_place;          /* cnt_x*cnt_y, bin k = y*cnt_x + x */
    const int64_t* non_place_unscaled; /* cnt_x*cnt_y */
    const int32_t* bin_x;              /* cnt_x + 1 bin edges along x (Bin::lx/ux) */
    const int32_t* bin_y;              /* cnt_y + 1 bin edges along y */
    const float* wx;                   /* cnt_x wavenumbers (nesterov.rs:351, x87 extended) */
    const float* wy;                   /* cnt_y wavenumbers (nesterov.rs:359) */
    int32_t group;                     /* -1: top level */
    /* updateBinsNonPlaceArea's target-independent halves (nesterov.rs:1956-2018), from which a new
     * target density rebuilds non_place and non_place_unscaled: per bin the fixed instances'
     * summed overlap areas, and the area of their union (-1 where no fixed instance touches the
     * bin). */
    const int64_t* non_place_raw;      /* cnt_x*cnt_y */
    const int64_t* non_place_union;    /* cnt_x*cnt_y */
} gpl_region;

typedef struct gpl_problem {
    int32_t num_cells;
    const int32_t* cell_box;      /* 4 per cell: lx, ly, ux, uy (GCell box) */
    const int32_t* cell_dbox;     /* 4 per cell: dlx, dly, dux, duy (density box) */
    const float* density_scale;   /* per cell */
    const uint8_t* cell_flags;    /* GPL_CELL_* */
    const int32_t* cell_inst;     /* placer instance of an instance cell, -1 for a filler */
    const int32_t* cell_pin_start;/* num_cells + 1: the cell's pins in its own pin order */
    const int32_t* cell_pins;
    int32_t num_pins;
    const int32_t* pin_cell;      /* -1: no cell (a port) */
    const int32_t* pin_offset;    /* 2 per pin: off_x, off_y */
    const int32_t* pin_xy;        /* 2 per pin: cx, cy as init leaves them */
    const int32_t* pin_net;
    int32_t num_nets;
    const int32_t* net_pin_start; /* num_nets + 1: the net's pins in net order */
    const int32_t* net_pins;
    const float* net_weight;      /* timing x custom weight (nesterov.rs:222), 1 */
    int32_t num_regions;
    const gpl_region* regions;
    gpl_opts opts;
    int32_t dbu;
    int32_t site_x;
    int32_t site_y;
    const int32_t* inst_box;      /* 4 per placer instance: the instance box (for placed_locations) */
    int32_t num_insts;
    int32_t die[4];               /* the die area (the RUDY grid's, RudyInputs.die) */
} gpl_problem;

/* Builds the Nesterov placer on the host (nesterov.rs:787 init_nesterov_place: cells, pins,
 * nets; per region fillers, perturbation, bin grid, fixed area, FFT wavenumbers, density
 * sizes).  The design keeps the arrays; gpl_design_problem returns a view of them valid until
 * the design is destroyed or re-initialised. */
int32_t gpl_init_nesterov_place(gpl_design* d, const gpl_opts* o);
int32_t gpl_design_problem(const gpl_design* d, gpl_problem* view);

/* ------------------------------------------------------------------------------------------ */
/* GPU context                                                                                */
/* ------------------------------------------------------------------------------------------ */

typedef struct gpl_ctx gpl_ctx;

/* Uploads a problem to device 0 (or the current device) and allocates the per-cell, per-pin,
 * per-net and per-bin state.  The problem's arrays are copied; they may be freed afterwards. */
int32_t gpl_ctx_create(const gpl_problem* p, gpl_ctx** out);
void gpl_ctx_destroy(gpl_ctx* c);
const char* gpl_last_error(const gpl_ctx* c);

/* Loop statistics (NesterovPlace's scalars plus per-region ones of region 0). */
typedef struct gpl_stats {
    int32_t iter;                  /* last iteration run */
    int32_t converged;
    int32_t status;                /* GPL_OK or a positive GPL code */
    int32_t total_trials;          /* back-tracking trials summed over iterations */
    int64_t hpwl;                  /* pin HPWL (nesterov.rs:2297) */
    float overflow;                /* average sum_overflow_unscaled */
    float overflow_scaled;         /* average sum_overflow */
    float density_penalty;         /* region 0 */
    float step_length;             /* region 0 */
    float wl_coef;                 /* wl_coef_x */
    float base_wl_coef;
    float reference_hpwl;
    float wl_grad_sum;             /* region 0 */
    float density_grad_sum;        /* region 0 */
    int32_t num_diverge_reverts;
    double init_seconds;           /* np_init wall time */
    double loop_seconds;           /* the loop's wall time */
    double first_below_target_seconds; /* np_init entry to the convergence iteration */
} gpl_stats;

/* nesterov.rs:4123: the reference HPWL from the design HPWL, then np_init (nesterov.rs:2674). */
int32_t gpl_init_nesterov_loop(gpl_ctx* c);
/* nesterov.rs:4440 resume_nesterov_place from iteration `start_iter`; returns the status and
 * writes the final iteration to *iter_out. */
int32_t gpl_run_nesterov_loop(gpl_ctx* c, int32_t start_iter, int32_t* iter_out);
/* nesterov.rs:4115: init, then the loop from 0. */
int32_t gpl_do_nesterov_place(gpl_ctx* c, int32_t* iter_out);
/* gpl_do_nesterov_place on each of n independent contexts (no reference counterpart: the crate
 * places one design at a time; for callers with many placements, e.g. seeds of one design).  The
 * contexts run in lockstep: each iteration's first back-tracking trial of every context is one
 * launch per kernel over all of them (contexts whose trials have the same kernels and grids, i.e.
 * the same shape, are grouped; the kernels' blocks take each context's arguments in turn), and
 * each context takes the rest of its iteration (its decisions, further trials, re-sorts) itself.
 * Each context's placement, rows and statistics are its own gpl_do_nesterov_place's, bit for bit.
 * iters (n entries, may be null) gets the final iterations.  Returns 0 or a negative API / HIP
 * error; a context's own outcome (a divergence's GPL code included) is in its gpl_get_stats.
 * Contexts with routability or profiling on run one after another instead. */
int32_t gpl_place_batch(gpl_ctx* const* ctxs, int32_t n, int32_t* iters);
/* nesterov.rs:4243: the loop again with a new target overflow and iteration limit. */
int32_t gpl_continue_nesterov_place(gpl_ctx* c, float overflow, int32_t max_iter, int32_t start_iter,
                                    int32_t* iter_out);
int32_t gpl_get_stats(gpl_ctx* c, gpl_stats* out);

/* The loop of gpl_run_nesterov_loop in slices, for callers that time or interleave it (no
 * reference counterpart; the three together are resume_nesterov_place, nesterov.rs:4440):
 * gpl_loop_begin does the per-call resets, gpl_loop_step runs at most n iterations (the device
 * work synchronised on return; *done is set once the loop ended: converged, diverged or out of
 * iterations), gpl_loop_end applies the max-iteration snapshot rule and writes the report. */
int32_t gpl_loop_begin(gpl_ctx* c, int32_t start_iter);
int32_t gpl_loop_step(gpl_ctx* c, int32_t n, int32_t* iter_out, int32_t* done);
int32_t gpl_loop_end(gpl_ctx* c, int32_t* iter_out);

/* Per-iteration history rows (iteration, overflow, HPWL, penalty, step, trials); returns the
 * number of rows written (at most `cap`). */
typedef struct gpl_iter_row {
    int32_t iter;
    int32_t trials;
    float overflow;
    float density_penalty;
    float step_length;
    float wl_coef;
    int64_t hpwl;
} gpl_iter_row;
int32_t gpl_get_history(gpl_ctx* c, gpl_iter_row* rows, int32_t cap);

/* The log lines written so far ("[INFO GPL-0001] ...", the progress table), joined with '\n'
 * into buf (truncated to cap bytes, NUL-terminated); returns the full length. */
int32_t gpl_get_log(gpl_ctx* c, char* buf, int32_t cap);

/* nesterov.rs:4099: each placed instance's lower-left (3 int32 per instance: placer instance
 * index, lx, ly); returns the number of instances written. */
int32_t gpl_placed_locations(gpl_ctx* c, int32_t* out, int32_t cap);

/* ------------------------------------------------------------------------------------------ */
/* Kernel-level entries (parity tests and the differentiable objective)                       */
/* ------------------------------------------------------------------------------------------ */

/* Which per-cell coordinate slot an entry reads or writes (Nb's vectors, nesterov.rs:480-493). */
#define GPL_SLOT_CUR_SLP 0
#define GPL_SLOT_NEXT_SLP 1
#define GPL_SLOT_PREV_SLP 2
#define GPL_SLOT_CUR 3
#define GPL_SLOT_NEXT 4

/* Host <-> device copies of per-cell float pairs (x, y interleaved) of one region. */
int32_t gpl_set_coordinates(gpl_ctx* c, int32_t region, int32_t slot, const float* xy);
int32_t gpl_get_coordinates(gpl_ctx* c, int32_t region, int32_t slot, float* xy);

/* updateGCellDensityCenterLocation + updateBinsGCellDensityArea (nesterov.rs:2268, 2218): the
 * region's cells' density centres from slot `slot`, the bins' integer areas, density and the
 * overflows.  Outputs (any may be NULL): inst/filler area per bin (int64, cnt_x*cnt_y), the
 * density per bin (float), and the two overflows as FP64 sums and as the integers the loop
 * keeps. */
int32_t gpl_density_scatter(gpl_ctx* c, int32_t region, int32_t slot, int64_t* inst_area,
                            int64_t* filler_area, float* density, double* overflow_f64);

/* updateDensityFieldBin (nesterov.rs:2279): the DCT Poisson solve of the region's current
 * density.  Outputs per bin (y-major, any may be NULL): potential, field x, field y. */
int32_t gpl_density_solve(gpl_ctx* c, int32_t region, float* phi, float* field_x, float* field_y);

/* Sets the density of a region directly (for solver tests), y-major. */
int32_t gpl_set_density(gpl_ctx* c, int32_t region, const float* density);

/* updateWireLengthForceWA (nesterov.rs:2325) at the current pin positions with the given WA
 * coefficients.  Outputs (any may be NULL): per net the 8 sums (min_x, x_min_x, max_x, x_max_x,
 * min_y, y_min_y, max_y, y_max_y) and box (lx, ly, ux, uy), and the pin HPWL. */
int32_t gpl_wa_forward(gpl_ctx* c, float wl_coef_x, float wl_coef_y, float* net_sums,
                       int32_t* net_box, int64_t* hpwl);

/* getWireLengthGradientWA for every cell of a region (nesterov.rs:2418) after gpl_wa_forward:
 * per cell x, y.  The pins' terms are those gpl_wa_forward computed (at the pin positions and
 * coefficients of that call), summed in each cell's pin order. */
int32_t gpl_wa_backward(gpl_ctx* c, int32_t region, float* wl_grad);

/* getDensityGradient for every cell of a region (nesterov.rs:2434) after gpl_density_solve. */
int32_t gpl_density_gather(gpl_ctx* c, int32_t region, float* den_grad);

/* updateGradients (nesterov.rs:2455) into a slot, at the current state (the WA terms of the last
 * gpl_wa_forward): per cell the WA and density gradients and the preconditioned sum; the FP64
 * sums |wl| and |den|. */
int32_t gpl_update_gradients(gpl_ctx* c, int32_t region, int32_t slot, float density_penalty,
                             float* wl, float* den, float* sum, double* abs_sums);

/* nesterovUpdateCoordinates' arithmetic (nesterov.rs:2718-2737) from the device state: next and
 * next look-ahead points for step `step` and momentum `coeff`. */
int32_t gpl_nesterov_step(gpl_ctx* c, int32_t region, float step, float coeff, float* next,
                          float* next_slp);

/* The differentiable objective (no reference counterpart: the reference evaluates only its
 * gradient, nesterov.rs:2392-2449): f(x) = sum_e w_e WA_e(x) + lambda * 1/2 sum_b rho_b phi_b over
 * region 0's cells at the continuous centres `xy` (2 floats per cell, problem order).  WA_e is the
 * weighted-average wirelength of nesterov.rs:2325-2388 with the exact exponential (expf) and
 * coefficient wl_coef; pins of region 0's cells sit at xy + offset, other pins at their initial
 * positions.  rho_b = (sum_i overlap_b(i) density_scale_i [x target for macros] + non_place_b) /
 * (area_b target) with each density box centred at its point (not truncated to integers), and
 * phi = K rho with K the region's Poisson solve (symmetric, so d/d rho of 1/2 rho^T K rho = phi).
 * forward: *value = wa + lambda * density, the two parts summed in FP64.  backward: df/dxy per
 * cell (float), the WA terms with the reference's double quotients.  Both overwrite region 0's
 * density, potential and fields (the loop recomputes them every trial). */
int32_t gpl_objective_forward(gpl_ctx* c, const float* xy, float wl_coef, float lambda,
                              double* value, double* wa_part, double* density_part);
int32_t gpl_objective_backward(gpl_ctx* c, const float* xy, float wl_coef, float lambda,
                               float* grad);

/* ---- routability's congestion measurement (route_base.rs) ---------------------------------- */
/* The RUDY grid (RudyInputs, route_base.rs:48-55, and makeGrid, 158-174): the die, tiles per side,
 * the tile size, the wire width (gpl_rudy_wire_width). */
typedef struct gpl_rudy_grid {
    int32_t lx, ly, ux, uy;
    int32_t cnt_x, cnt_y, tile_size, wire_width;
} gpl_rudy_grid;
/* RouteBaseVars' inflation fields (route_base.rs:12-46; defaults 0.95, 2, 3, 0.8, {1, 1, 0, 0}). */
typedef struct gpl_route_vars {
    float min_congestion_for_inflation, inflation_ratio_coef, max_inflation_ratio, ignore_edge_ratio;
    float rc_k[4];
} gpl_route_vars;
/* updateRudyAverage's figures (route_base.rs:493-536). */
typedef struct gpl_rudy_stats {
    double total_overflow;
    int32_t overflowed_tiles, actionable_tiles;
    double top[4]; /* the mean congestion of the top 0.5, 1, 2 and 5 % of the actionable tiles */
    float rc;      /* the weighted routing congestion (RouteBase::measure's cur_rc) */
} gpl_rudy_stats;
/* Rudy::new (route_base.rs:143-152): 1.25 / sum(1 / pitch), truncated; 100 with no pitch. */
int32_t gpl_rudy_wire_width(const int32_t* pitches, int32_t n);
/* Rudy::calculate (route_base.rs:183-253) on the device: every net's terminal box at the cells'
 * current centres deposits hpwl * wire_width / area, scaled by its share of each tile it strictly
 * overlaps, times 100, summed per tile in float in the nets' order (the reference's floats, bit for
 * bit); the blockage of each tile from (capacity, reduction) pairs (cap_red, 2 bytes per tile,
 * x-major; NULL: none).  demand and blockage (each cnt_x * cnt_y floats, x-major: x * cnt_y + y)
 * may be NULL; they stay on the device for gpl_rudy_tiles.  Overwrites the net records and the
 * pins' WA terms (the loop recomputes them every trial). */
int32_t gpl_rudy(gpl_ctx* c, const gpl_rudy_grid* g, const uint8_t* cap_red, float* demand, float* blockage);
/* calculateRudyTiles' tile half and updateRudyAverage (route_base.rs:443-536) on the last
 * gpl_rudy's demand: the tiles (y-major, iy * cnt_x + ix, origin (0, 0)) holding a standard
 * instance cell's density centre are actionable unless their blockage reaches
 * ignore_edge_ratio; an actionable tile's inflation ratio is min((congestion / min)^coef,
 * max) at congestion >= min, else 1; the statistics as the reference computes them (the top
 * shares over the congestions sorted in descending order).  inflation_ratio and actionable
 * (cnt_x * cnt_y each) may be NULL. */
int32_t gpl_rudy_tiles(gpl_ctx* c, const gpl_rudy_grid* g, const gpl_route_vars* v, float* inflation_ratio, uint8_t* actionable,
                       gpl_rudy_stats* stats);

/* A routability pass's cell inflation (RouteBase::routability's per-cell loop, nesterov.rs:3945-3960,
 * with set_size, 3754-3760): each standard instance cell of the region whose density centre lies in
 * a tile (y-major, origin (0, 0), the grid's tile size) of inflated ratio r > 1 grows to
 * round(dx sqrt r) x round(dy sqrt r), rebuilt from halves around its centre (an odd side loses
 * one).  sizes: the region's cell-box sizes (dx, dy per cell, the problem's order), updated in
 * place; *delta: the area added. */
int32_t gpl_route_inflate(gpl_ctx* c, int32_t region, const gpl_rudy_grid* g, const float* inflated_ratio, int32_t* sizes,
                          int64_t* delta);
/* updateDensitySize (nesterov.rs:2022-2036) of the region's cells at the given cell-box sizes (dx,
 * dy per cell, the problem's order): a side narrower than sqrt(2) bins (the reference's
 * extended-precision comparison) spreads to sqrt(2) bins with its density scaled down to match;
 * the density box rebuilt from halves around its centre; the preconditioner's cell area.  The
 * device cell model is updated in place. */
int32_t gpl_update_density_size(gpl_ctx* c, int32_t region, const int32_t* sizes);
/* The region's density boxes (4 per cell) and density scales, the problem's order (either may be
 * NULL). */
int32_t gpl_get_density_boxes(gpl_ctx* c, int32_t region, int32_t* dbox, float* density_scale);

/* Routability-driven placement's record (opts.routability; the loop's routability hooks of
 * resume_nesterov_place, nesterov.rs:4509-4567, and RouteBase::routability, 3985-4093): one row
 * per routability pass, and the totals. */
typedef struct gpl_route_pass {
    int32_t iter;             /* the iteration after which the pass ran */
    float rc;                 /* the measured weighted congestion (RouteBase::measure) */
    int32_t overflowed_tiles, actionable_tiles;
    int64_t inflation;        /* the area the pass's cell inflation added, over the regions */
    float target_density;     /* region 0's after the pass */
    int32_t fillers;          /* region 0's filler count after the pass */
    int32_t weighted_nets;    /* nets of weight above 1 after the pass */
    int32_t need, revert;     /* routability's (is_routability_need, is_revert_init_needed) */
} gpl_route_pass;
typedef struct gpl_route_stats {
    int32_t passes;           /* routability passes run */
    int32_t route_iters;      /* iterations counted as routability iterations (GPL-1017) */
    float final_rc;           /* the convergence's closing measurement (GPL-1005); 0 before */
    float min_rc;             /* the lowest congestion a pass kept (1e30 before the first pass) */
    int64_t total_inflation;  /* RouteBase::total_inflation */
    int64_t original_area;    /* the instances' area when the loop began (GPL-1011) */
} gpl_route_stats;
/* The routability record so far: the totals into *s (may be NULL), up to cap pass rows into
 * passes; returns the number of passes. */
int32_t gpl_get_route_stats(gpl_ctx* c, gpl_route_stats* s, gpl_route_pass* passes, int32_t cap);

/* The density map of gpl_density_scatter on arrays the caller keeps in device memory, with no
 * allocation (no reference counterpart: it serves callers that cannot call hipMalloc -- code under
 * HIP graph capture, tight serving loops, harnesses that forbid it).  The scatter of the density
 * boxes as given (updateBinsGCellDensityArea, nesterov.rs:2218-2250) into inst_area / filler_area
 * (int64 per bin, cleared first), the density per bin (2252-2263) and the two overflows
 * (overflow[0] scaled, overflow[1] unscaled; FP64 sums), all written on `stream` (a hipStream_t;
 * NULL: the null stream), asynchronously.  Its only other device memory is a fixed 33 KB array of
 * the library's (the overflow reduction's partials) per GPU, held until the process ends; calls
 * that overlap on one GPU race on it.  Limits: bin grids up to 4096 x 4096; dbox 16-byte aligned,
 * non_place_f 8-byte aligned; a larger grid or a misaligned array is refused with GPL_ERR_ARG and
 * a message on stderr.  Every pointer is a device pointer. */
typedef struct gpl_density_io {
    int32_t num_cells;
    const int32_t* dbox;          /* 4 per cell: dlx, dly, dux, duy (the density box) */
    const float* density_scale;   /* per cell */
    const uint8_t* cell_flags;    /* GPL_CELL_* per cell */
    int32_t lx, ly, ux, uy;       /* the bin grid (BinGrid, nesterov.rs:260) */
    int32_t cnt_x, cnt_y;
    double size_x, size_y;
    float target_density;
    const int32_t* bin_x;         /* cnt_x + 1 bin edges */
    const int32_t* bin_y;         /* cnt_y + 1 */
    const float* non_place_f;     /* 2 per bin: (float)non_place, (float)non_place_unscaled */
    int64_t* inst_area;           /* out: cnt_x * cnt_y, bin k = y * cnt_x + x */
    int64_t* filler_area;         /* out */
    float* density;               /* out */
    double* overflow;             /* out: 2 */
    void* stream;
} gpl_density_io;
int32_t gpl_density_scatter_noalloc(const gpl_density_io* io);

/* Fault injection for host-driven failure tests: after `after_iter` iterations the next
 * gradient sum reads NaN (GPL-0306) -- kind 1 -- or the next step length is NaN (GPL-0305) --
 * kind 2; kind 3 re-sorts the device cells after every iteration from `after_iter` on (the
 * storage order must not change the results); kind 4 (from now on, `after_iter` unused) runs
 * the kernels' general paths where they have fast ones (bin edges read from the region's tables,
 * the gradient's per-cell gathers of the pin terms), which must give the same results; kind 5:
 * in iteration `after_iter` the WA gradient sum is read as zero once (npUpdate*Gradient's
 * coefficient-halving recursion, nesterov.rs:2516-2522); kind 6: a divergence is reported in
 * iteration `after_iter` (a second call adds a second iteration), driving isDiverged's min-HPWL
 * snapshot revert and resume or stop (nesterov.rs:3009-3051).  Kind 0 clears every kind. */
int32_t gpl_debug_inject(gpl_ctx* c, int32_t kind, int32_t after_iter);

/* Device timing of the kernels while profiling (gpl_set_profiling; milliseconds summed per
 * kernel family and launch counts), for the roofline report: update_scatter, bin_density,
 * dct_solve, wa_forward, gradient, reduce; then "scatter_global_adds" (the scatter's atomic adds
 * into HBM in `ms`, the re-sorts in `launches`); then, once trials have run as graphs, the host's
 * time per phase of those trials: host_params (node arguments), host_launch (the launch call),
 * host_wait (the wait for the trial), host_between (from one trial's return to the next's
 * start), with the trial count in `launches`.  Returns the number of entries written. */
typedef struct gpl_kernel_time {
    char name[24];
    int32_t launches;
    double ms;
} gpl_kernel_time;
int32_t gpl_kernel_times(gpl_ctx* c, gpl_kernel_time* out, int32_t cap);

/* How a back-tracking trial is launched (no reference counterpart; the results are the same in
 * every mode): 0 (the default for one region) one HIP graph per trial, the bins and the Poisson
 * solve on a branch beside the WA kernels, the high-fanout nets on a third; 1 the same kernels
 * launched one by one
 * on those streams; 2 one by one on one stream; 3 one HIP graph per trial captured on one stream.
 * Device-driven trials (doBackTracking, nesterovUpdateStepLength, the phi / WA-coefficient
 * schedule and updateNextIter's bookkeeping, nesterov.rs:2750-2930, taken by the trial's last
 * kernel, so that trials follow each other without a host round trip; the host steps in where the
 * device halts on an event it handles): 4 as a HIP graph, 5 launched one by one on the side
 * streams, 6 on one stream.  gpl_kernel_times then reports "device_halts" (the halts, and the
 * host's time from each entry to the halt's state taken back) and "device_launched" (trials
 * launched, including those that found the device halted and did nothing).  7: grouped trials --
 * the trial's kernels recorded and launched as one launch per kernel variant and step over the
 * context's regions (every region's update at once, then every scatter, each row pass, the WA
 * kernels beside on the second stream, every gradient), the finishing reductions in order; the
 * default for a context of more than one region (mode 0 for one). */
int32_t gpl_set_launch_mode(gpl_ctx* c, int32_t mode);

/* The autotuner (no reference counterpart; north star: tile shapes picked by an autotuner the project
 * writes).  Per region it times the kernel variants whose results are bit-identical -- the scatter's
 * chunk of cells per block (256 or 512) and LDS sum width (32 bits where the bins allow, or 64), the
 * gradient's LDS staging of the pin terms (on or off) -- at gpl_init_nesterov_loop (before np_init,
 * outside its timing) and the scatter's again just before the shape's first re-sort (the storage
 * order at its most decayed, where the variants differ most), keeps the fastest, and caches the pick
 * per shape (cells, bin grid, pins) for the process.  gpl_set_tuning forces a region's
 * variant (chunk 256 or 512, table_bits 32 or 64, stage 0 or 1) and turns the tuner off; region -1
 * turns it on (chunk 1) or off (chunk 0).  gpl_get_tuning: per region 6 int32 -- chunk, table bits,
 * stage, source (0 default or forced, 1 measured, 2 cached), the scatter's and the gradient's
 * measured time in 1/100 us; returns the number of regions. */
int32_t gpl_set_tuning(gpl_ctx* c, int32_t region, int32_t chunk, int32_t table_bits, int32_t stage);
int32_t gpl_get_tuning(gpl_ctx* c, int32_t* out, int32_t cap);
int32_t gpl_set_profiling(gpl_ctx* c, int32_t on);
/* The device re-sort policy (no reference counterpart: the cells' storage order does not change
 * the placement): a re-sort when the scatter's global atomic adds per trial have grown by
 * `growth` (>= 1) over their level in the first iteration after the last re-sort, at least
 * `min_gap` (>= 1) iterations after it.  Defaults 50 and 1.5. */
int32_t gpl_set_resort_policy(gpl_ctx* c, int32_t min_gap, float growth);

#ifdef __cplusplus
}
#endif
#endif /* GPL_HIP_H */
