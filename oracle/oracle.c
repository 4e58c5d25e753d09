/*
 * TEST INFRASTRUCTURE -- the oracle.  A single-thread CPU restatement of the reference's
 * Nesterov placement loop (vyges-gpl 0.1.56, src/nesterov.rs), used as the parity checker of the
 * HIP path and as the CPU baseline the bench times.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product (gpl_amd/) never does.
 *
 * Input: the placer state init_nesterov_place leaves (nesterov.rs:787-933), as the C ABI's
 * gpl_problem holds it.  Every value keeps the reference's type and every expression its
 * evaluation order (float state, the double expressions where the reference has them, sums in
 * the reference's order), so a path that agrees with this agrees with the reference up to the
 * DCT's rounding (dct.c: this file's own FFT, not Ooura's butterflies).
 *
 * Besides the faithful loop it holds FP64 versions of the per-kernel arithmetic (bin density,
 * the Poisson solve, WA forward and backward, the density gather) for the tolerance checks.
 *
 * Pinning (tests/test_oracle_kat.py, test_oracle_dct.py, test_oracle_kernels.py, test_host.py):
 * the reference's known-answer tests (fast_exp nesterov.rs:4857,
 * distance 4893, wavenumbers 4697, reference HPWL 4813, divergence and snapshot rules
 * 4627-4669), brute-force sums on tiny inputs, closed forms (2-pin WA, single-mode Poisson),
 * invariants (translation invariance of WA, the scatter's area conservation) and scipy.fft.
 * Functions without such a check say "parity unpinned" in their comment.
 */
#define _POSIX_C_SOURCE 199309L
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../include/gpl_hip.h"

void orc_dct2_f32(int n, float* a);
void orc_dct3_f32(int n, float* a);
void orc_dst3_f32(int n, float* a);
void orc_ddct2d_fwd_f32(int n1, int n2, float* a);
void orc_ddct2d_inv_f32(int n1, int n2, float* a);
void orc_ddsct2d_inv_f32(int n1, int n2, float* a);
void orc_ddcst2d_inv_f32(int n1, int n2, float* a);
void orc_ddct2d_fwd_f64(int n1, int n2, double* a);
void orc_ddct2d_inv_f64(int n1, int n2, double* a);
void orc_ddsct2d_inv_f64(int n1, int n2, double* a);
void orc_ddcst2d_inv_f64(int n1, int n2, double* a);

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ---- small known-answer functions ---------------------------------------------------------- */

/* nesterov.rs:308 fastExp: (1 + x/1024)^1024 by ten squarings, in float. */
float orc_fast_exp(float x) {
    float e = 1.0f + x / 1024.0f;
    for (int i = 0; i < 10; ++i) e *= e;
    return e;
}

/* Host threads for the loop's independent per-cell, per-net, per-row and per-bin work (1 by
 * default: the reference is single-threaded).  Every sum whose order the reference fixes stays
 * sequential, and the bins' integer sums are exact in any order, so any count gives the same bits. */
static int g_threads = 1;
void orc_set_threads(int32_t n) { g_threads = n > 1 ? n : 1; }
int orc_threads(void) { return g_threads; }

/* Wide sums (a test switch, off by default): the loop's global float sums -- the step length's
 * distances, the gradients' |wl| and |den|, the bins' overflow -- accumulated in double instead of
 * the reference's float, the precision of the GPU path's FP64 reductions.  It separates what the
 * float sums' rounding contributes to a difference from what the arithmetic does. */
static int g_wide_sums = 0;
void orc_set_wide_sums(int32_t on) { g_wide_sums = on ? 1 : 0; }

/* nesterov.rs:319 getDistance: a float sum of squared differences in point order, then
 * sqrt(sum / (2 n)) in double.  a, b: n points, x and y interleaved. */
float orc_distance(const float* a, const float* b, int32_t n) {
    if (n == 0) return 0.0f;
    if (g_wide_sums) {
        double wide = 0.0;
        for (int32_t i = 0; i < n; ++i) {
            wide += (double)((a[2 * i] - b[2 * i]) * (a[2 * i] - b[2 * i]));
            wide += (double)((a[2 * i + 1] - b[2 * i + 1]) * (a[2 * i + 1] - b[2 * i + 1]));
        }
        return (float)sqrt(wide / (2.0 * (double)n));
    }
    float sum = 0.0f;
    for (int32_t i = 0; i < n; ++i) {
        sum += (a[2 * i] - b[2 * i]) * (a[2 * i] - b[2 * i]);
        sum += (a[2 * i + 1] - b[2 * i + 1]) * (a[2 * i + 1] - b[2 * i + 1]);
    }
    return (float)sqrt((double)sum / (2.0 * (double)n));
}

/* nesterov.rs:96 derived_reference_hpwl. */
float orc_derived_reference_hpwl(float user, int64_t design_hpwl) {
    if (user > 0.0f) return user;
    float f = 0.01f * (float)design_hpwl;
    return f > 446000000.0f ? f : 446000000.0f;
}

/* nesterov.rs:2532 updateWireLengthCoef's factor (times the base coefficient by the caller). */
float orc_wirelength_coef_factor(float overflow) {
    if (overflow > 1.0f) return (float)0.1;
    if ((double)overflow < 0.1) return 10.0f;
    return (float)(1.0 / pow(10.0, ((double)overflow - 0.1) * 20.0 / 9.0 - 1.0));
}

/* nesterov.rs:2827 getPhiCoef. */
float orc_phi_coef(float max_phi, float min_phi, float scaled_diff_hpwl) {
    float r = scaled_diff_hpwl < 0.0f ? max_phi : (float)((double)max_phi * pow((double)max_phi, -(double)scaled_diff_hpwl));
    return min_phi > r ? min_phi : r;
}

/* nesterov.rs:4406 diverge_fallback: 0 revert and resume, 1 revert and stop, 2 error. */
int32_t orc_diverge_fallback(int32_t saved, int32_t reverts, int32_t disable_revert) {
    if (!disable_revert && saved && reverts < 3) return 0;
    if (!disable_revert && reverts > 0) return 1;
    return 2;
}

/* nesterov.rs:4420 max_iter_takes_min_hpwl_snapshot. */
int32_t orc_max_iter_takes_snapshot(int32_t iter, int32_t max_iter, int32_t disable_revert, int32_t saved, int32_t reverts,
                                    int64_t snap_hpwl, float snapshot_overflow, float overflow, int64_t hpwl) {
    return iter >= max_iter && !disable_revert && (saved || reverts > 0) && snapshot_overflow <= overflow && snap_hpwl < hpwl;
}

/* ---- the Poisson solve (nesterov.rs:412-444 Fft::solve) ------------------------------------- */

/* Faithful: the density in x-major layout (a[x*ny + y]) in, potential and field out, all float;
 * the halving and scale in double as the reference writes them. */
void orc_fft_solve_f32(int32_t nx, int32_t ny, const float* wx, const float* wy, float* density, float* phi, float* field_x, float* field_y) {
    orc_ddct2d_fwd_f32(nx, ny, density);
    for (int i = 1; i < nx; ++i) density[(size_t)i * ny] = (float)((double)density[(size_t)i * ny] * 0.5);
    for (int j = 1; j < ny; ++j) density[j] = (float)((double)density[j] * 0.5);
    double scale = 4.0 / (double)nx / (double)ny;
    for (size_t k = 0; k < (size_t)nx * ny; ++k) density[k] = (float)((double)density[k] * scale);
    for (int i = 0; i < nx; ++i) {
        float wxi = wx[i], wx2 = wx[i] * wx[i];
        for (int j = 0; j < ny; ++j) {
            float wyj = wy[j], wy2 = wy[j] * wy[j];
            size_t k = (size_t)i * ny + j;
            float p = 0.0f, ex = 0.0f, ey = 0.0f;
            if (!(i == 0 && j == 0)) {
                p = density[k] / (wx2 + wy2);
                ex = p * wxi;
                ey = p * wyj;
            }
            phi[k] = p;
            field_x[k] = ex;
            field_y[k] = ey;
        }
    }
    orc_ddct2d_inv_f32(nx, ny, phi);
    orc_ddsct2d_inv_f32(nx, ny, field_x);
    orc_ddcst2d_inv_f32(nx, ny, field_y);
}

/* FP64: bins in y-major layout (k = y*nx + x) in and out, the same arithmetic in double with
 * the float wavenumbers widened. */
/* The solve of a double density (y-major), as orc_fft_solve_f64 below. */
void orc_fft_solve_f64d(int32_t nx, int32_t ny, const float* wx, const float* wy, const double* density_ym, double* phi_ym, double* fx_ym,
                        double* fy_ym) {
    size_t M = (size_t)nx * ny;
    double* a = (double*)malloc(sizeof(double) * M * 4);
    double *ph = a + M, *ex = a + 2 * M, *ey = a + 3 * M;
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) a[(size_t)x * ny + y] = density_ym[(size_t)y * nx + x];
    orc_ddct2d_fwd_f64(nx, ny, a);
    for (int i = 1; i < nx; ++i) a[(size_t)i * ny] *= 0.5;
    for (int j = 1; j < ny; ++j) a[j] *= 0.5;
    double scale = 4.0 / (double)nx / (double)ny;
    for (size_t k = 0; k < M; ++k) a[k] *= scale;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            size_t k = (size_t)i * ny + j;
            double w2 = (double)wx[i] * wx[i] + (double)wy[j] * wy[j];
            double p = (i == 0 && j == 0) ? 0.0 : a[k] / w2;
            ph[k] = p;
            ex[k] = p * wx[i];
            ey[k] = p * wy[j];
        }
    orc_ddct2d_inv_f64(nx, ny, ph);
    orc_ddsct2d_inv_f64(nx, ny, ex);
    orc_ddcst2d_inv_f64(nx, ny, ey);
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            size_t s = (size_t)x * ny + y, d = (size_t)y * nx + x;
            if (phi_ym) phi_ym[d] = ph[s];
            if (fx_ym) fx_ym[d] = ex[s];
            if (fy_ym) fy_ym[d] = ey[s];
        }
    free(a);
}

/* Fft::solve (nesterov.rs:412-444) in double: phi and the two fields of a float density. */
void orc_fft_solve_f64(int32_t nx, int32_t ny, const float* wx, const float* wy, const float* density_ym, double* phi_ym, double* fx_ym, double* fy_ym) {
    size_t M = (size_t)nx * ny;
    double* d = (double*)malloc(sizeof(double) * M);
    for (size_t k = 0; k < M; ++k) d[k] = (double)density_ym[k];
    orc_fft_solve_f64d(nx, ny, wx, wy, d, phi_ym, fx_ym, fy_ym);
    free(d);
}

/* ---- the placer state ----------------------------------------------------------------------- */

typedef struct {
    float x, y;
} fpt;

enum { S_CUR_SLP, S_CUR_SLP_WL, S_CUR_SLP_DEN, S_CUR_SLP_SUM, S_NEXT_SLP, S_NEXT_SLP_WL, S_NEXT_SLP_DEN, S_NEXT_SLP_SUM,
       S_PREV_SLP, S_PREV_SLP_WL, S_PREV_SLP_DEN, S_PREV_SLP_SUM, S_CUR, S_NEXT, S_SNAP, S_SNAP_SLP, S_SNAP_SLP_SUM,
       S_SNAP_PREV_SLP_SUM, S_COUNT };

typedef struct {
    gpl_region R;
    int32_t n, begin;
    int32_t* cells; /* the region's live cells (problem indices) in the reference's order: its
                     * instances, then its fillers (routability takes them from the back) */
    int32_t n_alloc, fillers;
    int64_t total_filler_area;
    /* the fillers routability took out since the last minimum, in removal order, each with its 18
     * per-cell values (RemovedFiller, nesterov.rs:552-571) */
    int32_t nrem;
    int32_t* rem_cell;
    fpt* rem_vals;
    int64_t *non_place_raw, *non_place_union;
    int64_t *non_place, *non_place_unscaled, *inst, *filler;
    int32_t *bin_x, *bin_y;
    float *wx, *wy;
    float *density, *field_x, *field_y, *phi, *fft_a, *fft_b, *fft_c, *fft_d;
    int64_t overflow, overflow_unscaled;
    fpt* v[S_COUNT];
    float snapshot_density_penalty, snapshot_step_length, step_length, density_penalty, wl_grad_sum, density_grad_sum;
    float sum_overflow, sum_overflow_unscaled, base_wl_coef_nb;
    int64_t prev_hpwl, prev_reported_hpwl;
    float prev_reported_overflow_unscaled, min_sum_overflow, hpwl_with_min_sum_overflow, coordi_distance, peak_coordi_distance;
    float max_phi_coef;
    int32_t max_phi_coef_changed, is_converged, is_diverged;
} orc_nb;

typedef struct orc_np {
    gpl_opts o;
    int32_t ncells, npins, nnets, nregions;
    /* cells */
    int32_t *lx, *ly, *ux, *uy, *dlx, *dly, *dux, *duy, *pin_start, *pins;
    int32_t *cnpins, *dxy; /* per cell: its pin count; its box's dx, dy (the preconditioner's inputs) */
    float* dscale;
    uint8_t* flags;
    /* pins */
    int32_t *pcx, *pcy, *poff_x, *poff_y, *pcell, *pnet;
    uint8_t* has;  /* bit 0 min_x, 1 max_x, 2 min_y, 3 max_y */
    float* pexp;   /* 4 per pin: min_x, max_x, min_y, max_y */
    /* nets */
    int32_t *net_start, *net_pins, *nbox;
    float *nw, *nsum; /* 8 per net */
    orc_nb* nbs;
    /* NesterovPlace */
    float wl_coef_x, wl_coef_y, base_wl_coef;
    int32_t recursion_wl, recursion_init_slp;
    float initial_prev_coordi_update_coef, reference_hpwl;
    int32_t max_iter;
    float average_overflow, average_overflow_unscaled;
    int64_t min_hpwl;
    int32_t is_min_hpwl, diverge_snapshot_iter;
    float diverge_snapshot_overflow;
    int32_t num_region_diverged, diverge_code;
    /* loop snapshots (nesterov.rs:4428) */
    int32_t diverge_saved, diverge_reverts;
    float diverge_coef_x, diverge_coef_y;
    int64_t diverge_hpwl;
    /* bookkeeping */
    int32_t trials, total_trials, iter_done, converged;
    double t_init, t_loop, t_converge;
    gpl_iter_row* hist;
    int32_t nhist, cap_hist;
    float* rrows; /* per iteration and region: iter, region, sum_overflow_unscaled, density_penalty, step, converged,
                   * wl_grad_sum, density_grad_sum */
    int32_t nrrows, cap_rrows;
    int32_t inject_nan_after;
    /* host-forced cases of the loop control (the GPU's gpl_debug_inject kinds 5 and 6): the WA
     * gradient sum read as zero once at iteration inject_zero_wl_at (the halving recursion), and a
     * divergence reported at iterations inject_diverge_at / inject_diverge_at2 (the min-HPWL
     * snapshot revert); -1: off */
    int32_t inject_zero_wl_at, inject_diverge_at, inject_diverge_at2, iter_now;
    /* a sensitivity probe: every field value of every solve scaled by (1 + field_noise u), u uniform
     * in [-1, 1) from a xorshift stream (0: off) -- rounding-level perturbations of the size the
     * GPU's transforms differ by, to measure how far the loop's outcome moves under them */
    float field_noise;
    uint64_t noise_state;
    /* routability (opts.routability): RouteBase (route_base.rs:279-326) and the placer's state */
    struct {
        int32_t on, die[4], cnt_x, cnt_y, tile, wire_width;
        float vars[8]; /* min congestion, coefficient, max ratio, ignore-edge ratio, rc_k[4] */
        int32_t revert_count, min_rc_violated;
        float min_rc;
        float *min_rc_target_density;
        int64_t *inflated_delta, *accumulated, *min_rc_inflated;
        int64_t original_movable_area;
        float *inflation_ratio, *inflated_ratio, *demand, *blockage; /* per tile */
        uint8_t* actionable;
        float *net_weight, *min_rc_net_weight, *net_demand; /* per net; 0: not in the map */
        int32_t *min_rc_dxy;                                /* per cell */
        int32_t need, settle_start, saved, attempts, route_iters, reverts, weighted;
        float coef_x, coef_y, a, final_rc;
        int32_t last_overflowed, last_actionable;
        gpl_route_pass* passes;
        int32_t npasses, cap_passes;
    } rt;
} orc_np;

static void* xcalloc(size_t n, size_t s) {
    void* p = calloc(n ? n : 1, s);
    if (!p) {
        fprintf(stderr, "oracle: out of memory\n");
        abort();
    }
    return p;
}

static void route_setup(orc_np* np, const gpl_problem* p);
static void route_free(orc_np* np);
static int try_routability_diverge_recovery(orc_np* np, float* cur_a);

orc_np* orc_create(const gpl_problem* p) {
    orc_np* np = (orc_np*)xcalloc(1, sizeof(orc_np));
    np->o = p->opts;
    int32_t N = p->num_cells, P = p->num_pins, E = p->num_nets;
    np->ncells = N;
    np->npins = P;
    np->nnets = E;
    np->lx = (int32_t*)xcalloc(N, 4);
    np->ly = (int32_t*)xcalloc(N, 4);
    np->ux = (int32_t*)xcalloc(N, 4);
    np->uy = (int32_t*)xcalloc(N, 4);
    np->dlx = (int32_t*)xcalloc(N, 4);
    np->dly = (int32_t*)xcalloc(N, 4);
    np->dux = (int32_t*)xcalloc(N, 4);
    np->duy = (int32_t*)xcalloc(N, 4);
    np->dscale = (float*)xcalloc(N, 4);
    np->flags = (uint8_t*)xcalloc(N, 1);
    np->pin_start = (int32_t*)xcalloc(N + 1, 4);
    np->pins = (int32_t*)xcalloc(p->cell_pin_start[N], 4);
    for (int32_t i = 0; i < N; ++i) {
        np->lx[i] = p->cell_box[4 * i];
        np->ly[i] = p->cell_box[4 * i + 1];
        np->ux[i] = p->cell_box[4 * i + 2];
        np->uy[i] = p->cell_box[4 * i + 3];
        np->dlx[i] = p->cell_dbox[4 * i];
        np->dly[i] = p->cell_dbox[4 * i + 1];
        np->dux[i] = p->cell_dbox[4 * i + 2];
        np->duy[i] = p->cell_dbox[4 * i + 3];
        np->dscale[i] = p->density_scale[i];
        np->flags[i] = p->cell_flags[i];
    }
    memcpy(np->pin_start, p->cell_pin_start, sizeof(int32_t) * (N + 1));
    np->cnpins = (int32_t*)xcalloc(N, 4);
    np->dxy = (int32_t*)xcalloc(2 * (size_t)N, 4);
    for (int32_t i = 0; i < N; ++i) {
        np->cnpins[i] = p->cell_pin_start[i + 1] - p->cell_pin_start[i];
        np->dxy[2 * i] = np->ux[i] - np->lx[i];
        np->dxy[2 * i + 1] = np->uy[i] - np->ly[i];
    }
    memcpy(np->pins, p->cell_pins, sizeof(int32_t) * p->cell_pin_start[N]);
    np->pcx = (int32_t*)xcalloc(P, 4);
    np->pcy = (int32_t*)xcalloc(P, 4);
    np->poff_x = (int32_t*)xcalloc(P, 4);
    np->poff_y = (int32_t*)xcalloc(P, 4);
    np->pcell = (int32_t*)xcalloc(P, 4);
    np->pnet = (int32_t*)xcalloc(P, 4);
    np->has = (uint8_t*)xcalloc(P, 1);
    np->pexp = (float*)xcalloc((size_t)P * 4, 4);
    for (int32_t q = 0; q < P; ++q) {
        np->pcx[q] = p->pin_xy[2 * q];
        np->pcy[q] = p->pin_xy[2 * q + 1];
        np->poff_x[q] = p->pin_offset[2 * q];
        np->poff_y[q] = p->pin_offset[2 * q + 1];
        np->pcell[q] = p->pin_cell[q];
        np->pnet[q] = p->pin_net[q];
    }
    np->net_start = (int32_t*)xcalloc(E + 1, 4);
    memcpy(np->net_start, p->net_pin_start, sizeof(int32_t) * (E + 1));
    np->net_pins = (int32_t*)xcalloc(p->net_pin_start[E], 4);
    memcpy(np->net_pins, p->net_pins, sizeof(int32_t) * p->net_pin_start[E]);
    np->nw = (float*)xcalloc(E, 4);
    memcpy(np->nw, p->net_weight, sizeof(float) * E);
    np->nbox = (int32_t*)xcalloc((size_t)E * 4, 4);
    np->nsum = (float*)xcalloc((size_t)E * 8, 4);
    np->nregions = p->num_regions;
    np->nbs = (orc_nb*)xcalloc(p->num_regions, sizeof(orc_nb));
    for (int32_t b = 0; b < p->num_regions; ++b) {
        orc_nb* nb = &np->nbs[b];
        nb->R = p->regions[b];
        nb->begin = nb->R.cell_begin;
        nb->n = nb->R.cell_end - nb->R.cell_begin;
        nb->n_alloc = nb->n;
        nb->cells = (int32_t*)xcalloc((size_t)nb->n, 4);
        for (int32_t i = 0; i < nb->n; ++i) nb->cells[i] = nb->begin + i;
        nb->fillers = nb->R.num_fillers;
        nb->total_filler_area = nb->R.total_filler_area;
        nb->rem_cell = (int32_t*)xcalloc((size_t)nb->R.num_fillers, 4);
        nb->rem_vals = (fpt*)xcalloc((size_t)nb->R.num_fillers * S_COUNT, sizeof(fpt));
        {
            size_t M0 = (size_t)nb->R.cnt_x * nb->R.cnt_y;
            nb->non_place_raw = (int64_t*)xcalloc(M0, 8);
            nb->non_place_union = (int64_t*)xcalloc(M0, 8);
            for (size_t k = 0; k < M0; ++k) {
                nb->non_place_raw[k] = nb->R.non_place_raw ? nb->R.non_place_raw[k] : 0;
                nb->non_place_union[k] = nb->R.non_place_union ? nb->R.non_place_union[k] : -1;
            }
        }
        size_t M = (size_t)nb->R.cnt_x * nb->R.cnt_y;
        nb->non_place = (int64_t*)xcalloc(M, 8);
        nb->non_place_unscaled = (int64_t*)xcalloc(M, 8);
        memcpy(nb->non_place, nb->R.non_place, 8 * M);
        memcpy(nb->non_place_unscaled, nb->R.non_place_unscaled, 8 * M);
        nb->inst = (int64_t*)xcalloc(M, 8);
        nb->filler = (int64_t*)xcalloc(M, 8);
        nb->bin_x = (int32_t*)xcalloc(nb->R.cnt_x + 1, 4);
        nb->bin_y = (int32_t*)xcalloc(nb->R.cnt_y + 1, 4);
        memcpy(nb->bin_x, nb->R.bin_x, 4 * (nb->R.cnt_x + 1));
        memcpy(nb->bin_y, nb->R.bin_y, 4 * (nb->R.cnt_y + 1));
        nb->wx = (float*)xcalloc(nb->R.cnt_x, 4);
        nb->wy = (float*)xcalloc(nb->R.cnt_y, 4);
        memcpy(nb->wx, nb->R.wx, 4 * nb->R.cnt_x);
        memcpy(nb->wy, nb->R.wy, 4 * nb->R.cnt_y);
        nb->density = (float*)xcalloc(M, 4);
        nb->field_x = (float*)xcalloc(M, 4);
        nb->field_y = (float*)xcalloc(M, 4);
        nb->phi = (float*)xcalloc(M, 4);
        nb->fft_a = (float*)xcalloc(M * 4, 4);
        nb->fft_b = nb->fft_a + M;
        nb->fft_c = nb->fft_a + 2 * M;
        nb->fft_d = nb->fft_a + 3 * M;
        for (int s = 0; s < S_COUNT; ++s) nb->v[s] = (fpt*)xcalloc(nb->n, sizeof(fpt));
        nb->max_phi_coef = np->o.max_phi_coef;
        nb->min_sum_overflow = 1e30f;
        nb->hpwl_with_min_sum_overflow = 1e30f;
    }
    np->initial_prev_coordi_update_coef = 100.0f;
    np->max_iter = np->o.max_iter;
    np->min_hpwl = INT64_MAX;
    np->inject_nan_after = -1;
    np->inject_zero_wl_at = -1;
    np->inject_diverge_at = -1;
    np->inject_diverge_at2 = -1;
    np->iter_now = -1;
    if (np->o.routability) route_setup(np, p);
    return np;
}

void orc_destroy(orc_np* np) {
    if (!np) return;
    for (int32_t b = 0; b < np->nregions; ++b) {
        orc_nb* nb = &np->nbs[b];
        free(nb->non_place);
        free(nb->non_place_unscaled);
        free(nb->inst);
        free(nb->filler);
        free(nb->bin_x);
        free(nb->bin_y);
        free(nb->wx);
        free(nb->wy);
        free(nb->density);
        free(nb->field_x);
        free(nb->field_y);
        free(nb->phi);
        free(nb->fft_a);
        free(nb->cells);
        free(nb->rem_cell);
        free(nb->rem_vals);
        free(nb->non_place_raw);
        free(nb->non_place_union);
        for (int s = 0; s < S_COUNT; ++s) free(nb->v[s]);
    }
    free(np->nbs);
    int32_t* ia[] = {np->lx, np->ly, np->ux, np->uy, np->dlx, np->dly, np->dux, np->duy, np->pin_start, np->pins, np->pcx, np->pcy,
                     np->poff_x, np->poff_y, np->pcell, np->pnet, np->net_start, np->net_pins, np->nbox, np->cnpins, np->dxy};
    for (size_t i = 0; i < sizeof ia / sizeof ia[0]; ++i) free(ia[i]);
    free(np->dscale);
    free(np->flags);
    free(np->has);
    free(np->pexp);
    free(np->nw);
    free(np->nsum);
    free(np->hist);
    free(np->rrows);
    route_free(np);
    free(np);
}

static int64_t nesterov_insts_area(const orc_nb* nb) {
    return nb->R.std_insts_area + (int64_t)roundf((float)nb->R.macro_insts_area * nb->R.target_density);
}

static int d_cx(const orc_np* np, int g) { return (np->dux[g] + np->dlx[g]) / 2; }
static int d_cy(const orc_np* np, int g) { return (np->duy[g] + np->dly[g]) / 2; }

/* nesterov.rs:2139 set_density_center with the cell's pins. */
static void set_density_center(orc_np* np, int g, int cx, int cy) {
    int hx = (np->dux[g] - np->dlx[g]) / 2, hy = (np->duy[g] - np->dly[g]) / 2;
    np->dlx[g] = cx - hx;
    np->dly[g] = cy - hy;
    np->dux[g] = cx + hx;
    np->duy[g] = cy + hy;
    int dcx = d_cx(np, g), dcy = d_cy(np, g);
    for (int k = np->pin_start[g]; k < np->pin_start[g + 1]; ++k) {
        int p = np->pins[k];
        np->pcx[p] = dcx + np->poff_x[p];
        np->pcy[p] = dcy + np->poff_y[p];
    }
}

/* nesterov.rs:2156 set_density_location. */
static void set_density_location(orc_np* np, int g, int lx, int ly) {
    np->dux[g] = lx + (np->dux[g] - np->dlx[g]);
    np->duy[g] = ly + (np->duy[g] - np->dly[g]);
    np->dlx[g] = lx;
    np->dly[g] = ly;
    int dcx = d_cx(np, g), dcy = d_cy(np, g);
    for (int k = np->pin_start[g]; k < np->pin_start[g + 1]; ++k) {
        int p = np->pins[k];
        np->pcx[p] = dcx + np->poff_x[p];
        np->pcy[p] = dcy + np->poff_y[p];
    }
}

/* nesterov.rs:2188-2213 getDensityCoordiLayoutInside{X,Y}: a centre c of a density box of integer
 * width `dd` kept inside [lo, hi] (the upper rule is applied last, so a box wider than the grid
 * ends at hi).  Pinned by hand-computed cases (tests/test_oracle_control.py). */
float orc_inside(float c, int32_t dd, int32_t lo, int32_t hi) {
    float half = (float)dd / 2.0f, adj = c;
    if (c - half < (float)lo) adj = (float)lo + half;
    if (c + half > (float)hi) adj = (float)hi - half;
    return adj;
}
static float inside_x(const orc_np* np, const orc_nb* nb, int g, float cx) { return orc_inside(cx, np->dux[g] - np->dlx[g], nb->R.lx, nb->R.ux); }
static float inside_y(const orc_np* np, const orc_nb* nb, int g, float cy) { return orc_inside(cy, np->duy[g] - np->dly[g], nb->R.ly, nb->R.uy); }

static void didx_x(const orc_nb* nb, int32_t lo, int32_t hi, int* x0, int* x1) {
    int a = (int)((double)(lo - nb->R.lx) / nb->R.size_x);
    int b = (int)ceil((double)(hi - nb->R.lx) / nb->R.size_x);
    *x0 = a > 0 ? a : 0;
    *x1 = b < nb->R.cnt_x ? b : nb->R.cnt_x;
}
static void didx_y(const orc_nb* nb, int32_t lo, int32_t hi, int* y0, int* y1) {
    int a = (int)((double)(lo - nb->R.ly) / nb->R.size_y);
    int b = (int)ceil((double)(hi - nb->R.ly) / nb->R.size_y);
    *y0 = a > 0 ? a : 0;
    *y1 = b < nb->R.cnt_y ? b : nb->R.cnt_y;
}

/* nesterov.rs:298 getOverlapDensityArea: a float product of the integer overlaps. */
static float overlap_density_area(int blx, int bly, int bux, int buy, int dlx, int dly, int dux, int duy) {
    int lx = blx > dlx ? blx : dlx, ly = bly > dly ? bly : dly;
    int ux = bux < dux ? bux : dux, uy = buy < duy ? buy : duy;
    if (lx >= ux || ly >= uy) return 0.0f;
    return (float)(ux - lx) * (float)(uy - ly);
}

/* nesterov.rs:2218 updateBinsGCellDensityArea. */
static void update_bins_density_area(orc_np* np, orc_nb* nb) {
    size_t M = (size_t)nb->R.cnt_x * nb->R.cnt_y;
    memset(nb->inst, 0, 8 * M);
    memset(nb->filler, 0, 8 * M);
    int cnt_x = nb->R.cnt_x;
    float td = nb->R.target_density;
    if (g_threads > 1) {
        /* the same integer areas from several threads, the adds atomic (exact in any order) */
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int i = 0; i < nb->n; ++i) {
            int g = nb->cells[i], x0, x1, y0, y1;
            didx_x(nb, np->dlx[g], np->dux[g], &x0, &x1);
            didx_y(nb, np->dly[g], np->duy[g], &y0, &y1);
            int inst = np->flags[g] & GPL_CELL_INSTANCE, macro = inst && (np->flags[g] & GPL_CELL_MACRO);
            for (int y = y0; y < y1; ++y)
                for (int x = x0; x < x1; ++x) {
                    float a = overlap_density_area(nb->bin_x[x], nb->bin_y[y], nb->bin_x[x + 1], nb->bin_y[y + 1], np->dlx[g], np->dly[g], np->dux[g], np->duy[g]) * np->dscale[g];
                    size_t k = (size_t)y * cnt_x + x;
                    if (inst && macro) a *= td;
                    int64_t* dst = inst ? &nb->inst[k] : &nb->filler[k];
#pragma omp atomic
                    *dst += (int64_t)a;
                }
        }
    } else {
        for (int i = 0; i < nb->n; ++i) {
            int g = nb->cells[i], x0, x1, y0, y1;
            didx_x(nb, np->dlx[g], np->dux[g], &x0, &x1);
            didx_y(nb, np->dly[g], np->duy[g], &y0, &y1);
            int inst = np->flags[g] & GPL_CELL_INSTANCE, macro = inst && (np->flags[g] & GPL_CELL_MACRO);
            for (int y = y0; y < y1; ++y)
                for (int x = x0; x < x1; ++x) {
                    float a = overlap_density_area(nb->bin_x[x], nb->bin_y[y], nb->bin_x[x + 1], nb->bin_y[y + 1], np->dlx[g], np->dly[g], np->dux[g], np->duy[g]) * np->dscale[g];
                    size_t k = (size_t)y * cnt_x + x;
                    if (inst) {
                        if (macro) a *= td;
                        nb->inst[k] += (int64_t)a;
                    } else {
                        nb->filler[k] += (int64_t)a;
                    }
                }
        }
    }
    nb->overflow = 0;
    nb->overflow_unscaled = 0;
    double wo = 0.0, wou = 0.0;
    for (int y = 0; y < nb->R.cnt_y; ++y)
        for (int x = 0; x < cnt_x; ++x) {
            size_t k = (size_t)y * cnt_x + x;
            int64_t area = (int64_t)(nb->bin_x[x + 1] - nb->bin_x[x]) * (int64_t)(nb->bin_y[y + 1] - nb->bin_y[y]);
            float scaled = (float)area * td;
            nb->density[k] = ((float)nb->inst[k] + (float)nb->filler[k] + (float)nb->non_place[k]) / scaled;
            float o = (float)nb->inst[k] + (float)nb->non_place[k] - scaled;
            o = o > 0.0f ? o : 0.0f;
            nb->overflow = (int64_t)((float)nb->overflow + o);
            float ou = (float)nb->inst[k] + (float)nb->non_place_unscaled[k] - scaled;
            ou = ou > 0.0f ? ou : 0.0f;
            nb->overflow_unscaled = (int64_t)((float)nb->overflow_unscaled + ou);
            if (g_wide_sums) {
                wo += (double)o;
                wou += (double)ou;
            }
        }
    if (g_wide_sums) {
        nb->overflow = (int64_t)wo;
        nb->overflow_unscaled = (int64_t)wou;
    }
}

/* nesterov.rs:2268 updateGCellDensityCenterLocation then the bins. */
static void update_center_location(orc_np* np, orc_nb* nb, const fpt* c) {
    if (g_threads > 1) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int i = 0; i < nb->n; ++i) set_density_center(np, nb->cells[i], (int)c[i].x, (int)c[i].y);
    } else {
        for (int i = 0; i < nb->n; ++i) set_density_center(np, nb->cells[i], (int)c[i].x, (int)c[i].y);
    }
    update_bins_density_area(np, nb);
}

/* nesterov.rs:2279 updateDensityFieldBin. */
static float noise_factor(orc_np* np) {
    uint64_t x = np->noise_state;
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    np->noise_state = x;
    return 1.0f + np->field_noise * ((float)(x & 0xffffu) / 32768.0f - 1.0f);
}

static void density_to_xmajor(orc_nb* nb, float* a, int y) {
    const int nx = nb->R.cnt_x, ny = nb->R.cnt_y;
    for (int x = 0; x < nx; ++x) a[(size_t)x * ny + y] = nb->density[(size_t)y * nx + x];
}
static void fields_from_xmajor(orc_np* np, orc_nb* nb, const float* ph, const float* fx, const float* fy, int y) {
    const int nx = nb->R.cnt_x, ny = nb->R.cnt_y;
    for (int x = 0; x < nx; ++x) {
        size_t s = (size_t)x * ny + y, d = (size_t)y * nx + x;
        nb->field_x[d] = fx[s];
        nb->field_y[d] = fy[s];
        nb->phi[d] = ph[s];
        if (np && np->field_noise != 0.0f) {
            nb->field_x[d] *= noise_factor(np);
            nb->field_y[d] *= noise_factor(np);
        }
    }
}

static void update_density_field_bin(orc_np* np, orc_nb* nb) {
    int ny = nb->R.cnt_y;
    float *a = nb->fft_a, *ph = nb->fft_b, *fx = nb->fft_c, *fy = nb->fft_d;
    /* (the noise probe draws its factors in order: one thread) */
    const int par = g_threads > 1 && !(np && np->field_noise != 0.0f);
    if (par) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int y = 0; y < ny; ++y) density_to_xmajor(nb, a, y);
    } else {
        for (int y = 0; y < ny; ++y) density_to_xmajor(nb, a, y);
    }
    orc_fft_solve_f32(nb->R.cnt_x, ny, nb->wx, nb->wy, a, ph, fx, fy);
    if (par) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int y = 0; y < ny; ++y) fields_from_xmajor(np, nb, ph, fx, fy, y);
    } else {
        for (int y = 0; y < ny; ++y) fields_from_xmajor(np, nb, ph, fx, fy, y);
    }
}

/* Turns the sensitivity probe on (eps > 0, stream seeded with `seed`) or off (eps = 0). */
void orc_set_field_noise(orc_np* np, float eps, uint64_t seed) {
    np->field_noise = eps;
    np->noise_state = seed ? seed : 88172645463325252ull;
}

/* nesterov.rs:2307 GNet::updateBox. */
static void update_box(orc_np* np, int n) {
    int lx = INT32_MAX, ly = INT32_MAX, ux = INT32_MIN, uy = INT32_MIN;
    for (int k = np->net_start[n]; k < np->net_start[n + 1]; ++k) {
        int p = np->net_pins[k];
        if (np->pcx[p] < lx) lx = np->pcx[p];
        if (np->pcy[p] < ly) ly = np->pcy[p];
        if (np->pcx[p] > ux) ux = np->pcx[p];
        if (np->pcy[p] > uy) uy = np->pcy[p];
    }
    np->nbox[4 * n] = lx;
    np->nbox[4 * n + 1] = ly;
    np->nbox[4 * n + 2] = ux;
    np->nbox[4 * n + 3] = uy;
}

/* nesterov.rs:2297 getHpwl. */
static inline int64_t net_hpwl(orc_np* np, int n) {
    update_box(np, n);
    const int32_t* b = &np->nbox[4 * n];
    return b[2] < b[0] ? 0 : ((int64_t)b[2] - b[0]) + ((int64_t)b[3] - b[1]);
}

int64_t orc_np_hpwl(orc_np* np) {
    int64_t total = 0;
    if (g_threads > 1) {
#pragma omp parallel for num_threads(g_threads) schedule(static) reduction(+ : total)
        for (int n = 0; n < np->nnets; ++n) total += net_hpwl(np, n);
    } else {
        for (int n = 0; n < np->nnets; ++n) total += net_hpwl(np, n);
    }
    return total;
}

/* nesterov.rs:2325 updateWireLengthForceWA: one net (its box, its pins' exponentials and sums). */
static inline void wa_net(orc_np* np, int n, float cx, float cy) {
    float* s = &np->nsum[8 * n];
    for (int q = 0; q < 8; ++q) s[q] = 0.0f;
    update_box(np, n);
    int nlx = np->nbox[4 * n], nly = np->nbox[4 * n + 1], nux = np->nbox[4 * n + 2], nuy = np->nbox[4 * n + 3];
    for (int k = np->net_start[n]; k < np->net_start[n + 1]; ++k) {
        int p = np->net_pins[k], px = np->pcx[p], py = np->pcy[p];
        float emin_x = (float)(nlx - px) * cx, emax_x = (float)(px - nux) * cx;
        float emin_y = (float)(nly - py) * cy, emax_y = (float)(py - nuy) * cy;
        if (emin_x > -300.0f) {
            float v = orc_fast_exp(emin_x);
            np->has[p] |= 1;
            np->pexp[4 * p] = v;
            s[0] += v;
            s[1] += (float)px * v;
        }
        if (emax_x > -300.0f) {
            float v = orc_fast_exp(emax_x);
            np->has[p] |= 2;
            np->pexp[4 * p + 1] = v;
            s[2] += v;
            s[3] += (float)px * v;
        }
        if (emin_y > -300.0f) {
            float v = orc_fast_exp(emin_y);
            np->has[p] |= 4;
            np->pexp[4 * p + 2] = v;
            s[4] += v;
            s[5] += (float)py * v;
        }
        if (emax_y > -300.0f) {
            float v = orc_fast_exp(emax_y);
            np->has[p] |= 8;
            np->pexp[4 * p + 3] = v;
            s[6] += v;
            s[7] += (float)py * v;
        }
    }
}

/* nesterov.rs:2325 updateWireLengthForceWA (nets are independent: a pin is on one net). */
static void update_wirelength_force_wa(orc_np* np) {
    float cx = np->wl_coef_x, cy = np->wl_coef_y;
    memset(np->has, 0, np->npins);
    memset(np->pexp, 0, sizeof(float) * 4 * (size_t)np->npins);
    if (g_threads > 1) {
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 256)
        for (int n = 0; n < np->nnets; ++n) wa_net(np, n, cx, cy);
    } else {
        for (int n = 0; n < np->nnets; ++n) wa_net(np, n, cx, cy);
    }
}

/* nesterov.rs:2392 getWireLengthGradientPinWA. */
static fpt wl_gradient_pin(const orc_np* np, int p, int n) {
    const float* g = &np->nsum[8 * n];
    const float* q = &np->pexp[4 * p];
    float cx = np->wl_coef_x, cy = np->wl_coef_y;
    float gmin_x = 0.0f, gmax_x = 0.0f, gmin_y = 0.0f, gmax_y = 0.0f;
    float qcx = (float)np->pcx[p], qcy = (float)np->pcy[p];
    if (np->has[p] & 1) {
        double num = (double)g[0] * ((double)q[0] * (1.0 - (double)(cx * qcx))) + (double)(cx * q[0] * g[1]);
        gmin_x = (float)(num / (double)(g[0] * g[0]));
    }
    if (np->has[p] & 2) {
        double num = (double)g[2] * ((double)q[1] * (1.0 + (double)(cx * qcx))) - (double)(cx * q[1] * g[3]);
        gmax_x = (float)(num / (double)(g[2] * g[2]));
    }
    if (np->has[p] & 4) {
        double num = (double)g[4] * ((double)q[2] * (1.0 - (double)(cy * qcy))) + (double)(cy * q[2] * g[5]);
        gmin_y = (float)(num / (double)(g[4] * g[4]));
    }
    if (np->has[p] & 8) {
        double num = (double)g[6] * ((double)q[3] * (1.0 + (double)(cy * qcy))) - (double)(cy * q[3] * g[7]);
        gmax_y = (float)(num / (double)(g[6] * g[6]));
    }
    fpt r = {gmin_x - gmax_x, gmin_y - gmax_y};
    return r;
}

/* nesterov.rs:2418 getWireLengthGradientWA. */
static fpt wl_gradient(const orc_np* np, int g) {
    fpt s = {0.0f, 0.0f};
    for (int k = np->pin_start[g]; k < np->pin_start[g + 1]; ++k) {
        int p = np->pins[k], n = np->pnet[p];
        fpt t = wl_gradient_pin(np, p, n);
        float w = np->nw[n];
        t.x *= w;
        t.y *= w;
        s.x += t.x;
        s.y += t.y;
    }
    return s;
}

/* nesterov.rs:2434 getDensityGradient: x outer, y inner. */
static fpt density_gradient(const orc_np* np, const orc_nb* nb, int g) {
    int x0, x1, y0, y1;
    didx_x(nb, np->dlx[g], np->dux[g], &x0, &x1);
    didx_y(nb, np->dly[g], np->duy[g], &y0, &y1);
    fpt f = {0.0f, 0.0f};
    for (int x = x0; x < x1; ++x)
        for (int y = y0; y < y1; ++y) {
            size_t k = (size_t)y * nb->R.cnt_x + x;
            float a = overlap_density_area(nb->bin_x[x], nb->bin_y[y], nb->bin_x[x + 1], nb->bin_y[y + 1], np->dlx[g], np->dly[g], np->dux[g], np->duy[g]) * np->dscale[g];
            f.x += a * nb->field_x[k];
            f.y += a * nb->field_y[k];
        }
    return f;
}

enum { SLOT_PREV, SLOT_CUR, SLOT_NEXT };

/* nesterov.rs:2469-2489, updateGradients' arithmetic on given gradients: per cell the
 * preconditioned sum (wl + p den) / max(pins + p (dx dy), 1) -- the floor taken as Rust's
 * f32::max, which returns 1 for a NaN preconditioner -- and the float sums of |wl| and |den| in
 * cell order, x before y.  wl, den, sum: n points (x, y interleaved); dxy: 2 ints per cell (the
 * cell box's dx, dy).  Pinned by hand-computed cases and the convex-quadratic run
 * (tests/test_oracle_control.py). */
void orc_precondition(int32_t n, const float* wl, const float* den, const int32_t* npins, const int32_t* dxy, float p, float* sum,
                      float* wl_abs, float* den_abs) {
    for (int32_t i = 0; i < n; ++i) {
        float sx = wl[2 * i] + p * den[2 * i];
        float sy = wl[2 * i + 1] + p * den[2 * i + 1];
        float wl_pre = (float)npins[i];
        float area = (float)dxy[2 * i] * (float)dxy[2 * i + 1];
        float pre = wl_pre + p * area;
        if (!(pre >= 1.0f)) pre = 1.0f;
        sum[2 * i] = sx / pre;
        sum[2 * i + 1] = sy / pre;
    }
    if (g_wide_sums) {
        double wa = 0.0, wd = 0.0;
        for (int32_t i = 0; i < n; ++i) {
            wa += (double)fabsf(wl[2 * i]) + (double)fabsf(wl[2 * i + 1]);
            wd += (double)fabsf(den[2 * i]) + (double)fabsf(den[2 * i + 1]);
        }
        *wl_abs = (float)wa;
        *den_abs = (float)wd;
        return;
    }
    float a = 0.0f, d = 0.0f;
    for (int32_t i = 0; i < n; ++i) {
        a += fabsf(wl[2 * i]);
        a += fabsf(wl[2 * i + 1]);
        d += fabsf(den[2 * i]);
        d += fabsf(den[2 * i + 1]);
    }
    *wl_abs = a;
    *den_abs = d;
}

/* nesterov.rs:2455 updateGradients into a slot. */
static void update_gradients(orc_np* np, orc_nb* nb, int slot) {
    if (nb->is_converged) return;
    int base = slot == SLOT_PREV ? S_PREV_SLP : slot == SLOT_CUR ? S_CUR_SLP : S_NEXT_SLP;
    fpt *wl = nb->v[base + 1], *den = nb->v[base + 2], *sum = nb->v[base + 3];
    if (g_threads > 1) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int i = 0; i < nb->n; ++i) {
            wl[i] = wl_gradient(np, nb->cells[i]);
            den[i] = density_gradient(np, nb, nb->cells[i]);
        }
    } else {
        for (int i = 0; i < nb->n; ++i) wl[i] = wl_gradient(np, nb->cells[i]);
        for (int i = 0; i < nb->n; ++i) den[i] = density_gradient(np, nb, nb->cells[i]);
    }
    int32_t* pins = (int32_t*)xcalloc((size_t)nb->n, 4);
    int32_t* dxy = (int32_t*)xcalloc(2 * (size_t)nb->n, 4);
    for (int i = 0; i < nb->n; ++i) {
        pins[i] = np->cnpins[nb->cells[i]];
        dxy[2 * i] = np->dxy[2 * (size_t)nb->cells[i]];
        dxy[2 * i + 1] = np->dxy[2 * (size_t)nb->cells[i] + 1];
    }
    orc_precondition(nb->n, (const float*)wl, (const float*)den, pins, dxy, nb->density_penalty, (float*)sum, &nb->wl_grad_sum,
                     &nb->density_grad_sum);
    free(pins);
    free(dxy);
}

static int bad_f(float v) { return isnan(v) || isinf(v); }

/* nesterov.rs:2512 npUpdate*Gradient. */
static void np_update_gradient(orc_np* np, int b, int slot) {
    orc_nb* nb = &np->nbs[b];
    update_gradients(np, nb, slot);
    if (np->inject_nan_after >= 0 && np->iter_done >= np->inject_nan_after && slot == SLOT_NEXT) nb->wl_grad_sum = NAN;
    if (np->inject_zero_wl_at >= 0 && np->iter_now == np->inject_zero_wl_at && slot == SLOT_NEXT) {
        nb->wl_grad_sum = 0.0f;
        np->inject_zero_wl_at = -1;
    }
    if (nb->wl_grad_sum == 0.0f && np->recursion_wl < 10) {
        np->wl_coef_x = (float)((double)np->wl_coef_x * 0.5);
        np->wl_coef_y = (float)((double)np->wl_coef_y * 0.5);
        np->base_wl_coef = (float)((double)np->base_wl_coef * 0.5);
        update_wirelength_force_wa(np);
        np->recursion_wl += 1;
        np_update_gradient(np, b, slot);
        return;
    }
    if (bad_f(nb->wl_grad_sum) || bad_f(nb->density_grad_sum)) {
        np->diverge_code = 306;
        np->num_region_diverged = 1;
    }
}

/* nesterov.rs:2532 updateWireLengthCoef. */
static void update_wirelength_coef(orc_np* np, float overflow) {
    float c = orc_wirelength_coef_factor(overflow);
    np->wl_coef_x = c * np->base_wl_coef;
    np->wl_coef_y = c * np->base_wl_coef;
}

/* nesterov.rs:2546 initDensity1. */
static void init_density1(orc_np* np, int b) {
    orc_nb* nb = &np->nbs[b];
    for (int i = 0; i < nb->n; ++i) {
        int g = nb->cells[i];
        /* updateDensityCoordiLayoutInside (nesterov.rs:2173) */
        float tlx = (float)np->dlx[g] > (float)nb->R.lx ? (float)np->dlx[g] : (float)nb->R.lx;
        float tly = (float)np->dly[g] > (float)nb->R.ly ? (float)np->dly[g] : (float)nb->R.ly;
        int ddx = np->dux[g] - np->dlx[g], ddy = np->duy[g] - np->dly[g];
        if (tlx + (float)ddx > (float)nb->R.ux) tlx = (float)(nb->R.ux - ddx);
        if (tly + (float)ddy > (float)nb->R.uy) tly = (float)(nb->R.uy - ddy);
        set_density_location(np, g, (int)tlx, (int)tly);
        fpt p = {(float)d_cx(np, g), (float)d_cy(np, g)};
        nb->v[S_CUR_SLP][i] = p;
        nb->v[S_PREV_SLP][i] = p;
        nb->v[S_CUR][i] = p;
    }
    update_center_location(np, nb, nb->v[S_CUR_SLP]);
    nb->prev_hpwl = orc_np_hpwl(np);
    update_density_field_bin(np, nb);
    nb->base_wl_coef_nb = (float)((double)np->o.init_wirelength_coef / ((double)(float)(nb->R.size_x + nb->R.size_y) * 0.5));
    float nia = (float)nesterov_insts_area(nb);
    nb->sum_overflow = (float)nb->overflow / nia;
    nb->sum_overflow_unscaled = (float)nb->overflow_unscaled / nia;
}

/* nesterov.rs:2583 updateInitialPrevSLPCoordi. */
static void update_initial_prev_slp(orc_np* np, int b) {
    orc_nb* nb = &np->nbs[b];
    float c = np->initial_prev_coordi_update_coef;
    for (int i = 0; i < nb->n; ++i) {
        int g = nb->cells[i];
        if (np->flags[g] & GPL_CELL_LOCKED) {
            nb->v[S_PREV_SLP][i] = nb->v[S_CUR_SLP][i];
            continue;
        }
        float px = nb->v[S_CUR_SLP][i].x - c * nb->v[S_CUR_SLP_SUM][i].x;
        float py = nb->v[S_CUR_SLP][i].y - c * nb->v[S_CUR_SLP_SUM][i].y;
        nb->v[S_PREV_SLP][i].x = inside_x(np, nb, g, px);
        nb->v[S_PREV_SLP][i].y = inside_y(np, nb, g, py);
    }
}

/* nesterov.rs:2608-2611 initDensity2's first penalty: (WA sum / density sum) x the initial factor;
 * with a zero WA sum the penalty is left as it is (`cur`).  Pinned by hand-computed cases
 * (tests/test_oracle_control.py). */
float orc_init_penalty(float wl_grad_sum, float density_grad_sum, float cur, float init_density_penalty) {
    if (wl_grad_sum != 0.0f) return (wl_grad_sum / density_grad_sum) * init_density_penalty;
    return cur;
}

/* nesterov.rs:2615-2617 initDensity2's first step: the RMS distance of the two points over that of
 * their preconditioned gradients.  Pinned as orc_step_length. */
float orc_init_step(int32_t n, const float* prev_slp, const float* cur_slp, const float* prev_sum, const float* cur_sum, float* coordi_distance) {
    *coordi_distance = orc_distance(prev_slp, cur_slp, n);
    float grad = orc_distance(prev_sum, cur_sum, n);
    return *coordi_distance / grad;
}

/* nesterov.rs:2600 initDensity2. */
static float init_density2(orc_np* np, int b) {
    orc_nb* nb = &np->nbs[b];
    if (nb->wl_grad_sum == 0.0f) {
        nb->density_penalty = np->o.init_density_penalty;
        update_gradients(np, nb, SLOT_PREV);
    }
    nb->density_penalty = orc_init_penalty(nb->wl_grad_sum, nb->density_grad_sum, nb->density_penalty, np->o.init_density_penalty);
    float nia = (float)nesterov_insts_area(nb);
    nb->sum_overflow = (float)nb->overflow / nia;
    nb->sum_overflow_unscaled = (float)nb->overflow_unscaled / nia;
    nb->step_length = orc_init_step(nb->n, (const float*)nb->v[S_PREV_SLP], (const float*)nb->v[S_CUR_SLP], (const float*)nb->v[S_PREV_SLP_SUM],
                                    (const float*)nb->v[S_CUR_SLP_SUM], &nb->coordi_distance);
    return nb->step_length;
}

/* nesterov.rs:2674 NesterovPlace::init. */
static int32_t np_init(orc_np* np) {
    float total_sum_overflow = 0.0f, total_base = 0.0f;
    for (int b = 0; b < np->nregions; ++b) {
        init_density1(np, b);
        total_sum_overflow += np->nbs[b].sum_overflow;
        total_base += np->nbs[b].base_wl_coef_nb;
    }
    float n = (float)np->nregions;
    np->average_overflow = total_sum_overflow / n;
    np->base_wl_coef = total_base / n;
    update_wirelength_coef(np, np->average_overflow);
    update_wirelength_force_wa(np);
    for (int b = 0; b < np->nregions; ++b) {
        orc_nb* nb = &np->nbs[b];
        np_update_gradient(np, b, SLOT_CUR);
        update_initial_prev_slp(np, b);
        update_center_location(np, nb, nb->v[S_PREV_SLP]);
        update_density_field_bin(np, nb);
    }
    update_wirelength_force_wa(np);
    for (int b = 0; b < np->nregions; ++b) np_update_gradient(np, b, SLOT_PREV);
    for (int b = 0; b < np->nregions; ++b) {
        float step = init_density2(np, b);
        if (bad_f(step) && np->recursion_init_slp < 10) {
            np->initial_prev_coordi_update_coef *= 10.0f;
            np->recursion_init_slp += 1;
            return np_init(np);
        }
        if (bad_f(step)) return 304;
    }
    return 0;
}

/* nesterov.rs:4123 init_nesterov_loop. */
int32_t orc_init(orc_np* np) {
    double t0 = now_s();
    int64_t design_hpwl = orc_np_hpwl(np);
    np->reference_hpwl = orc_derived_reference_hpwl(np->o.reference_hpwl, design_hpwl);
    int32_t rc = np_init(np);
    np->t_init = now_s() - t0;
    if (rc) {
        np->num_region_diverged = 1;
        np->diverge_code = rc;
    }
    return rc;
}

/* nesterov.rs:2718-2746 nesterovUpdateCoordinates' per-cell arithmetic: next = slp + s sum and
 * next_slp = next + coeff (next - cur), each kept inside the grid box (lx, ly, ux, uy) by the
 * cell's density box size dd (2 ints per cell); a cell with a nonzero `locked` entry keeps cur and
 * cur_slp.  Points are n (x, y) pairs.  Pinned by hand-computed cases and the convex-quadratic
 * run (tests/test_oracle_control.py). */
void orc_nesterov_points(int32_t n, const float* cur_slp, const float* sum, const float* cur, const uint8_t* locked, const int32_t* dd,
                         const int32_t* box, float s, float coeff, float* next, float* next_slp) {
    for (int32_t k = 0; k < n; ++k) {
        if (locked[k]) {
            next[2 * k] = cur[2 * k];
            next[2 * k + 1] = cur[2 * k + 1];
            next_slp[2 * k] = cur_slp[2 * k];
            next_slp[2 * k + 1] = cur_slp[2 * k + 1];
            continue;
        }
        float nx = cur_slp[2 * k] + s * sum[2 * k];
        float ny = cur_slp[2 * k + 1] + s * sum[2 * k + 1];
        float sx = nx + coeff * (nx - cur[2 * k]);
        float sy = ny + coeff * (ny - cur[2 * k + 1]);
        next[2 * k] = orc_inside(nx, dd[2 * k], box[0], box[2]);
        next[2 * k + 1] = orc_inside(ny, dd[2 * k + 1], box[1], box[3]);
        next_slp[2 * k] = orc_inside(sx, dd[2 * k], box[0], box[2]);
        next_slp[2 * k + 1] = orc_inside(sy, dd[2 * k + 1], box[1], box[3]);
    }
}

/* nesterov.rs:2718 nesterovUpdateCoordinates. */
static void nesterov_update_coordinates(orc_np* np, int b, float coeff) {
    orc_nb* nb = &np->nbs[b];
    if (nb->is_converged) return;
    uint8_t* locked = (uint8_t*)xcalloc(nb->n, 1);
    int32_t* dd = (int32_t*)xcalloc(2 * (size_t)nb->n, 4);
    for (int k = 0; k < nb->n; ++k) {
        int g = nb->cells[k];
        locked[k] = (np->flags[g] & GPL_CELL_LOCKED) ? 1 : 0;
        dd[2 * k] = np->dux[g] - np->dlx[g];
        dd[2 * k + 1] = np->duy[g] - np->dly[g];
    }
    const int32_t box[4] = {nb->R.lx, nb->R.ly, nb->R.ux, nb->R.uy};
    orc_nesterov_points(nb->n, (const float*)nb->v[S_CUR_SLP], (const float*)nb->v[S_CUR_SLP_SUM], (const float*)nb->v[S_CUR], locked, dd,
                        box, nb->step_length, coeff, (float*)nb->v[S_NEXT], (float*)nb->v[S_NEXT_SLP]);
    free(locked);
    free(dd);
    update_center_location(np, nb, nb->v[S_NEXT_SLP]);
    update_density_field_bin(np, nb);
}

/* nesterov.rs:2750-2772 nesterovUpdateStepLength on given points: the new step is the RMS distance
 * of the look-ahead points over that of their preconditioned gradients (getDistance, 319); a NaN
 * or infinite one sets *diverged and is refused; one above 0.95 of the old (compared in double) is
 * taken but not accepted; one under 0.01 becomes 0.01, not accepted; else accepted.  Returns 1
 * when accepted; *step is the old step in and the new out.  Pinned by hand-computed cases, the
 * isotropic-quadratic closed form (the step equals the preconditioner over the curvature) and the
 * convex-quadratic run (tests/test_oracle_control.py). */
int32_t orc_step_length(int32_t n, const float* cur_slp, const float* next_slp, const float* cur_sum, const float* next_sum, float* step,
                        float* coordi_distance, int32_t* diverged) {
    *coordi_distance = orc_distance(cur_slp, next_slp, n);
    float grad = orc_distance(cur_sum, next_sum, n);
    float nw = *coordi_distance / grad;
    if (bad_f(nw)) {
        *diverged = 1;
        return 0;
    }
    if ((double)nw > (double)*step * 0.95) {
        *step = nw;
        return 0;
    }
    if ((double)nw < 0.01) {
        *step = (float)0.01;
        return 0;
    }
    *step = nw;
    return 1;
}

/* nesterov.rs:2750 nesterovUpdateStepLength. */
static int nesterov_update_step_length(orc_np* np, orc_nb* nb) {
    if (nb->is_converged) return 1;
    return orc_step_length(nb->n, (const float*)nb->v[S_CUR_SLP], (const float*)nb->v[S_NEXT_SLP], (const float*)nb->v[S_CUR_SLP_SUM],
                           (const float*)nb->v[S_NEXT_SLP_SUM], &nb->step_length, &nb->coordi_distance, &nb->is_diverged);
}

/* nesterov.rs:2776 doBackTracking. */
static void do_back_tracking(orc_np* np, float coeff) {
    np->trials = 0;
    for (int nbt = 0; nbt < 10; ++nbt) {
        np->trials++;
        for (int b = 0; b < np->nregions; ++b) nesterov_update_coordinates(np, b, coeff);
        update_wirelength_force_wa(np);
        np->num_region_diverged = 0;
        for (int b = 0; b < np->nregions; ++b) {
            np_update_gradient(np, b, SLOT_NEXT);
            np->num_region_diverged += np->nbs[b].is_diverged;
        }
        if (np->num_region_diverged > 0) {
            np->diverge_code = 306;
            break;
        }
        int ok = 0;
        np->num_region_diverged = 0;
        for (int b = 0; b < np->nregions; ++b) {
            ok += nesterov_update_step_length(np, &np->nbs[b]);
            np->num_region_diverged += np->nbs[b].is_diverged;
        }
        if (np->num_region_diverged > 0) {
            np->diverge_code = 305;
            break;
        }
        if (ok != np->nregions) break;
    }
}

/* nesterov.rs:2812 nesterovAdjustPhi. */
static void nesterov_adjust_phi(orc_nb* nb) {
    if (nb->is_converged) return;
    if (!nb->max_phi_coef_changed && nb->sum_overflow_unscaled < 0.35f) {
        nb->max_phi_coef_changed = 1;
        nb->max_phi_coef = (float)((double)nb->max_phi_coef * 0.99);
    }
    if (nb->max_phi_coef <= 1.0f) nb->max_phi_coef = 1.01f;
}

static void swap_slots(orc_nb* nb, int a, int b) {
    fpt* t = nb->v[a];
    nb->v[a] = nb->v[b];
    nb->v[b] = t;
}

/* nesterov.rs:2837 NesterovBase::updateNextIter. */
static void nb_update_next_iter(orc_np* np, int b, int iter) {
    orc_nb* nb = &np->nbs[b];
    if (nb->is_converged) return;
    for (int q = 0; q < 4; ++q) swap_slots(nb, S_PREV_SLP + q, S_CUR_SLP + q);
    for (int k = 0; k < nb->n; ++k) {
        int g = nb->cells[k];
        if ((np->flags[g] & GPL_CELL_INSTANCE) && (np->flags[g] & GPL_CELL_LOCKED)) {
            for (int q = 0; q < 4; ++q) nb->v[S_NEXT_SLP + q][k] = nb->v[S_CUR_SLP + q][k];
            nb->v[S_NEXT][k] = nb->v[S_CUR][k];
        }
    }
    for (int q = 0; q < 4; ++q) swap_slots(nb, S_CUR_SLP + q, S_NEXT_SLP + q);
    swap_slots(nb, S_CUR, S_NEXT);
    float fraction = (float)iter / (float)np->max_iter;
    float a = (float)nesterov_insts_area(nb), c = fraction * (float)nb->R.non_place_insts_area * 0.05f;
    float denom = a > c ? a : c;
    nb->sum_overflow = (float)nb->overflow / denom;
    nb->sum_overflow_unscaled = (float)nb->overflow_unscaled / denom;
    int64_t hpwl = orc_np_hpwl(np);
    if (iter == 0 || iter % 10 == 0) {
        nb->prev_reported_hpwl = hpwl;
        nb->prev_reported_overflow_unscaled = nb->sum_overflow_unscaled;
    }
    float scaled = (float)(hpwl - nb->prev_hpwl) / np->reference_hpwl;
    float phi = orc_phi_coef(nb->max_phi_coef, np->o.min_phi_coef, scaled);
    nb->density_penalty *= phi;
    nb->prev_hpwl = hpwl;
    if (nb->coordi_distance > nb->peak_coordi_distance) nb->peak_coordi_distance = nb->coordi_distance;
    if (iter > 50 && nb->min_sum_overflow > nb->sum_overflow_unscaled) {
        nb->min_sum_overflow = nb->sum_overflow_unscaled;
        nb->hpwl_with_min_sum_overflow = (float)nb->prev_hpwl;
    }
}

/* nesterov.rs:2907 NesterovPlace::updateNextIter. */
static void np_update_next_iter(orc_np* np, int iter) {
    float total = 0.0f, total_unscaled = 0.0f;
    for (int b = 0; b < np->nregions; ++b) {
        nb_update_next_iter(np, b, iter);
        total += np->nbs[b].sum_overflow;
        total_unscaled += np->nbs[b].sum_overflow_unscaled;
    }
    float n = (float)np->nregions;
    np->average_overflow = total / n;
    np->average_overflow_unscaled = total_unscaled / n;
    update_wirelength_coef(np, np->average_overflow);
    if (!np->o.disable_revert_if_diverge) {
        int64_t hpwl = orc_np_hpwl(np);
        if (hpwl < np->min_hpwl && (double)np->average_overflow_unscaled <= 0.25) {
            np->min_hpwl = hpwl;
            np->diverge_snapshot_overflow = np->average_overflow_unscaled;
            np->diverge_snapshot_iter = iter + 1;
            np->is_min_hpwl = 1;
        } else {
            np->is_min_hpwl = 0;
        }
    }
}

/* nesterov.rs:2933 saveSnapshot / 2950 revertToSnapshot. */
static void save_snapshot(orc_nb* nb) {
    if (nb->is_converged) return;
    memcpy(nb->v[S_SNAP], nb->v[S_CUR], sizeof(fpt) * nb->n);
    memcpy(nb->v[S_SNAP_SLP], nb->v[S_CUR_SLP], sizeof(fpt) * nb->n);
    memcpy(nb->v[S_SNAP_SLP_SUM], nb->v[S_CUR_SLP_SUM], sizeof(fpt) * nb->n);
    memcpy(nb->v[S_SNAP_PREV_SLP_SUM], nb->v[S_PREV_SLP_SUM], sizeof(fpt) * nb->n);
    nb->snapshot_density_penalty = nb->density_penalty;
    nb->snapshot_step_length = nb->step_length;
}

static void revert_to_snapshot(orc_np* np, orc_nb* nb) {
    if (nb->is_converged) return;
    memcpy(nb->v[S_CUR], nb->v[S_SNAP], sizeof(fpt) * nb->n);
    memcpy(nb->v[S_CUR_SLP], nb->v[S_SNAP_SLP], sizeof(fpt) * nb->n);
    memcpy(nb->v[S_CUR_SLP_SUM], nb->v[S_SNAP_SLP_SUM], sizeof(fpt) * nb->n);
    memcpy(nb->v[S_PREV_SLP_SUM], nb->v[S_SNAP_PREV_SLP_SUM], sizeof(fpt) * nb->n);
    nb->density_penalty = nb->snapshot_density_penalty;
    nb->step_length = nb->snapshot_step_length;
    update_center_location(np, nb, nb->v[S_CUR]);
    update_density_field_bin(np, nb);
    nb->is_diverged = 0;
    nb->prev_reported_hpwl = 0;
    nb->prev_reported_overflow_unscaled = 0.0f;
}

/* nesterov.rs:2976-2991 checkDivergence's two rules: bit 0 (GPL-0323) the overflow is under 0.2,
 * at least 0.02 above its minimum, and the HPWL more than 1.2x the HPWL at that minimum; bit 1
 * (GPL-0324) since the last progress row the overflow rose by at least 0.05 and the HPWL by at
 * least 25 %.  Pinned by hand-computed cases on both sides of every bound
 * (tests/test_oracle_control.py). */
int32_t orc_check_divergence(float sum_overflow_unscaled, float min_sum_overflow, float hpwl_with_min_sum_overflow, int64_t prev_hpwl,
                             float prev_reported_overflow_unscaled, int64_t prev_reported_hpwl) {
    int32_t r = 0;
    if (sum_overflow_unscaled < 0.2f && sum_overflow_unscaled - min_sum_overflow >= 0.02f && hpwl_with_min_sum_overflow * 1.2f < (float)prev_hpwl)
        r |= 1;
    if (prev_reported_overflow_unscaled > 0.0f && prev_reported_hpwl > 0) {
        float change = sum_overflow_unscaled - prev_reported_overflow_unscaled;
        float increase = (float)(prev_hpwl - prev_reported_hpwl) / (float)prev_reported_hpwl;
        if (change >= 0.05f && increase >= 0.25f) r |= 2;
    }
    return r;
}

/* nesterov.rs:2976 checkDivergence. */
static int check_divergence(orc_np* np, orc_nb* nb) {
    if (orc_check_divergence(nb->sum_overflow_unscaled, nb->min_sum_overflow, nb->hpwl_with_min_sum_overflow, nb->prev_hpwl,
                             nb->prev_reported_overflow_unscaled, nb->prev_reported_hpwl))
        nb->is_diverged = 1;
    if (np->inject_diverge_at >= 0 && np->iter_now == np->inject_diverge_at) nb->is_diverged = 1;
    if (np->inject_diverge_at2 >= 0 && np->iter_now == np->inject_diverge_at2) nb->is_diverged = 1;
    return nb->is_diverged;
}

static void revert_to_diverge_snapshot(orc_np* np) {
    np->wl_coef_x = np->diverge_coef_x;
    np->wl_coef_y = np->diverge_coef_y;
    update_wirelength_force_wa(np);
    for (int b = 0; b < np->nregions; ++b) {
        revert_to_snapshot(np, &np->nbs[b]);
        np->nbs[b].min_sum_overflow = 1e30f;
        np->nbs[b].hpwl_with_min_sum_overflow = 1e30f;
    }
}

/* nesterov.rs:3009 isDiverged: 0 none, 1 resume, 2 stop.  The min-HPWL snapshot only once
 * routability is off; a divergence tries the routability snapshot first. */
static int np_is_diverged(orc_np* np, float* cur_a) {
    np->num_region_diverged = 0;
    for (int b = 0; b < np->nregions; ++b) np->num_region_diverged += check_divergence(np, &np->nbs[b]);
    const int routability_off = !np->rt.on || !np->rt.need;
    if (np->num_region_diverged == 0 && np->is_min_hpwl && !np->o.disable_revert_if_diverge && routability_off) {
        np->diverge_coef_x = np->wl_coef_x;
        np->diverge_coef_y = np->wl_coef_y;
        np->diverge_hpwl = np->min_hpwl;
        for (int b = 0; b < np->nregions; ++b) save_snapshot(&np->nbs[b]);
        np->diverge_saved = 1;
    }
    if (np->num_region_diverged == 0) return 0;
    if (!np->o.disable_revert_if_diverge && np->rt.on && try_routability_diverge_recovery(np, cur_a)) {
        np->num_region_diverged = 0;
        return 1;
    }
    switch (orc_diverge_fallback(np->diverge_saved, np->diverge_reverts, np->o.disable_revert_if_diverge)) {
    case 0:
        np->diverge_reverts += 1;
        revert_to_diverge_snapshot(np);
        *cur_a = 1.0f;
        np->diverge_saved = 0;
        np->num_region_diverged = 0;
        return 1;
    case 1:
        revert_to_diverge_snapshot(np);
        np->num_region_diverged = 0;
        return 2;
    default:
        np->diverge_code = 307;
        return 2;
    }
}

/* ---- routability-driven placement (opts.routability) ------------------------------------------
 * RouteBase::routability (nesterov.rs:3985-4093) with its helpers (cutFillerCells 3826-3870,
 * restoreRemovedFillers 3873-3891, revertToMinCongestion 3896-3918, applyNetWeights 3922-3937,
 * scaleInflationToBudget 3946-3978, setTargetDensity 3793-3801, updateAreas 3804-3822), RouteBase's
 * measurement (route_base.rs:443-536 through route.c) and net weighting (381-426), and the loop's
 * routability hooks (nesterov.rs:4509-4567, 3009-3090, 3132-3140).  The reference's cell order
 * throughout: a removed filler leaves from the back of its region's cells with its per-cell values
 * and comes back at the end in removal order. */
int32_t orc_rudy_wire_width(const int32_t* pitches, int32_t n);
void orc_rudy(const int32_t die[4], int32_t cnt_x, int32_t cnt_y, int32_t tile, int32_t wire_width, int32_t nnets, const int32_t* boxes,
              const uint8_t* cap_red, float* demand, float* blockage);
void orc_rudy_tiles(int32_t cnt_x, int32_t cnt_y, int32_t tile, const float* demand, const float* blockage, int32_t ncentres,
                    const int32_t* centres, const float* vars, float* inflation_ratio, uint8_t* actionable, double* stats);
void orc_set_size(const int32_t* box, int32_t dx, int32_t dy, int32_t* out);
void orc_rudy_rect(const int32_t die[4], int32_t cnt_x, int32_t cnt_y, int32_t tile, int32_t x, int32_t y, int32_t* r);

static int is_std(const orc_np* np, int g) { return (np->flags[g] & GPL_CELL_INSTANCE) && !(np->flags[g] & GPL_CELL_MACRO); }

static void route_setup(orc_np* np, const gpl_problem* p) {
    const orc_nb* nb0 = &np->nbs[0];
    int32_t* die = np->rt.die;
    memcpy(die, p->die, sizeof(int32_t) * 4);
    if (die[2] <= die[0] || die[3] <= die[1]) { /* region 0's box when the problem names no die */
        die[0] = nb0->R.lx;
        die[1] = nb0->R.ly;
        die[2] = nb0->R.ux;
        die[3] = nb0->R.uy;
    }
    const int32_t gcells = np->o.route_gcells > 0 ? np->o.route_gcells : 64;
    int32_t tile = (die[2] - die[0]) / gcells;
    if (tile < 1) tile = 1;
    int32_t pitch = tile / (np->o.route_tracks > 0 ? np->o.route_tracks : 8);
    if (pitch < 1) pitch = 1;
    const int32_t pitches[2] = {pitch, pitch};
    np->rt.on = 1;
    np->rt.cnt_x = gcells;
    np->rt.cnt_y = (die[3] - die[1]) / tile > 0 ? (die[3] - die[1]) / tile : 1;
    np->rt.tile = tile;
    np->rt.wire_width = orc_rudy_wire_width(pitches, 2);
    const float v[8] = {np->o.routability_min_congestion, np->o.routability_inflation_ratio_coef, np->o.routability_max_inflation_ratio, 0.8f,
                        np->o.routability_rc_k[0], np->o.routability_rc_k[1], np->o.routability_rc_k[2], np->o.routability_rc_k[3]};
    memcpy(np->rt.vars, v, sizeof v);
    np->rt.min_rc = 1e30f;
    int R = np->nregions;
    np->rt.min_rc_target_density = (float*)xcalloc(R, 4);
    np->rt.inflated_delta = (int64_t*)xcalloc(R, 8);
    np->rt.accumulated = (int64_t*)xcalloc(R, 8);
    np->rt.min_rc_inflated = (int64_t*)xcalloc(R, 8);
    for (int b = 0; b < R; ++b) np->rt.original_movable_area += nesterov_insts_area(&np->nbs[b]);
    size_t T = (size_t)np->rt.cnt_x * np->rt.cnt_y;
    np->rt.inflation_ratio = (float*)xcalloc(T, 4);
    np->rt.inflated_ratio = (float*)xcalloc(T, 4);
    np->rt.demand = (float*)xcalloc(T, 4);
    np->rt.blockage = (float*)xcalloc(T, 4);
    np->rt.actionable = (uint8_t*)xcalloc(T, 1);
    np->rt.net_weight = (float*)xcalloc(np->nnets, 4);
    np->rt.min_rc_net_weight = (float*)xcalloc(np->nnets, 4);
    np->rt.net_demand = (float*)xcalloc(np->nnets, 4);
    np->rt.min_rc_dxy = (int32_t*)xcalloc(2 * (size_t)np->ncells, 4);
    np->rt.need = 1;
    np->rt.settle_start = -1;
    np->rt.a = 1.0f;
}

static void route_free(orc_np* np) {
    void* f[] = {np->rt.min_rc_target_density, np->rt.inflated_delta, np->rt.accumulated, np->rt.min_rc_inflated, np->rt.inflation_ratio,
                 np->rt.inflated_ratio, np->rt.demand, np->rt.blockage, np->rt.actionable, np->rt.net_weight, np->rt.min_rc_net_weight,
                 np->rt.net_demand, np->rt.min_rc_dxy, np->rt.passes};
    for (size_t i = 0; i < sizeof f / sizeof f[0]; ++i) free(f[i]);
}

/* updateBinsNonPlaceArea's last steps (nesterov.rs:1990-2017) at target density td from the
 * target-independent halves: a bin's non_place is its fixed instances' summed overlap (raw) times
 * td, capped at its union times td times 1.10 where a fixed instance touches it (uni >= 0), and
 * non_place_unscaled its union times td -- float products truncated to int64 as the reference's
 * `as i64`.  Pinned by the construction's arrays at the initial target density (every bin of the
 * 1M-cell design's macro map) and hand-computed bins (tests/test_oracle_route.py). */
void orc_non_place_at(int32_t m, const int64_t* raw, const int64_t* uni, float td, int64_t* non_place, int64_t* non_place_unscaled) {
    for (int32_t k = 0; k < m; ++k) {
        const int64_t r = raw[k], u = uni[k];
        int64_t v = r != 0 ? (int64_t)((float)r * td) : 0, vu = 0;
        if (u >= 0) {
            vu = (int64_t)((float)u * td);
            const int64_t cap = (int64_t)((float)u * td * 1.10f);
            if (v > cap) v = cap;
        }
        non_place[k] = v;
        non_place_unscaled[k] = vu;
    }
}

/* setTargetDensity: the region's and its bins' target density, the fixed area again at it. */
static void set_target_density(orc_nb* nb, float td) {
    nb->R.target_density = td;
    orc_non_place_at(nb->R.cnt_x * nb->R.cnt_y, nb->non_place_raw, nb->non_place_union, td, nb->non_place, nb->non_place_unscaled);
}

/* updateAreas: the standard and macro instance areas from the cells' sizes now. */
static void update_areas(orc_np* np, orc_nb* nb) {
    int64_t st = 0, mac = 0;
    for (int i = 0; i < nb->n; ++i) {
        const int g = nb->cells[i];
        if (!(np->flags[g] & GPL_CELL_INSTANCE)) continue;
        const int64_t a = (int64_t)(np->ux[g] - np->lx[g]) * (int64_t)(np->uy[g] - np->ly[g]);
        if (np->flags[g] & GPL_CELL_MACRO)
            mac += a;
        else
            st += a;
    }
    nb->R.std_insts_area = st;
    nb->R.macro_insts_area = mac;
}

/* updateDensitySize (nesterov.rs:2022-2036) of the region's cells, the density box rebuilt around
 * its centre (set_density_size, 171-177). */
static void update_density_size_nb(orc_np* np, orc_nb* nb) {
    const long double s2 = 1.414213562373095048801L;
    const long double thr_x = s2 * (long double)nb->R.size_x, thr_y = s2 * (long double)nb->R.size_y;
    const float size_x = (float)(s2 * (long double)(float)nb->R.size_x), size_y = (float)(s2 * (long double)(float)nb->R.size_y);
    for (int i = 0; i < nb->n; ++i) {
        const int g = nb->cells[i];
        const int32_t dx = np->ux[g] - np->lx[g], dy = np->uy[g] - np->ly[g];
        float sx = 1.0f, sy = 1.0f, dsx = (float)dx, dsy = (float)dy;
        if ((long double)(uint64_t)dx < thr_x) {
            sx = (float)dx / (float)thr_x;
            dsx = size_x;
        }
        if ((long double)(uint64_t)dy < thr_y) {
            sy = (float)dy / (float)thr_y;
            dsy = size_y;
        }
        const int32_t ddx = (int32_t)dsx, ddy = (int32_t)dsy, cx = d_cx(np, g), cy = d_cy(np, g);
        np->dlx[g] = cx - ddx / 2;
        np->dly[g] = cy - ddy / 2;
        np->dux[g] = cx + ddx / 2;
        np->duy[g] = cy + ddy / 2;
        np->dscale[g] = sx * sy;
    }
}

/* cutFillerCells: whole fillers from the back of the region's cells for the inflation's area
 * (each with its values, swapAndPop of the last), the target density raised for the rest. */
/* cutFillerCells' arithmetic (nesterov.rs:3834-3866): how many whole fillers of area `single` an
 * inflation of `area` takes out of `fillers`, and the target density when more than a filler's
 * area is left over -- (instance area + removed + fillers left + remainder) / white space, in
 * float; *td is left as it is otherwise.  Returns the count.  Pinned by hand-computed cases
 * (tests/test_oracle_route.py). */
int32_t orc_cut_fillers_plan(int64_t area, int64_t single, int32_t fillers, int64_t insts_area, int64_t white_space, float* td) {
    if (area < 0) return 0;
    int64_t max_remove = single > 0 ? area / single : 0;
    if (max_remove > fillers) max_remove = fillers;
    const int64_t left = single * ((int64_t)fillers - max_remove), removed_area = single * max_remove, remaining = area - removed_area;
    if (remaining > single) *td = (float)(insts_area + removed_area + left + remaining) / (float)white_space;
    return (int32_t)max_remove;
}

static void cut_filler_cells(orc_np* np, orc_nb* nb, int64_t area) {
    if (area < 0) return; /* GPL-0313 */
    const int64_t single = (int64_t)nb->R.filler_dx * nb->R.filler_dy;
    float td = nb->R.target_density;
    const int64_t max_remove = orc_cut_fillers_plan(area, single, nb->fillers, nesterov_insts_area(nb), nb->R.white_space_area, &td);
    int64_t available = single * nb->fillers, removed = 0;
    for (int i = nb->n - 1; i >= 0 && removed < max_remove; --i) {
        const int g = nb->cells[i];
        if (np->flags[g] & GPL_CELL_INSTANCE) continue;
        /* swap_pop(i): the last cell moves to i (i is the last here: the back holds the fillers) */
        const int last = nb->n - 1;
        nb->rem_cell[nb->nrem] = g;
        for (int s = 0; s < S_COUNT; ++s) nb->rem_vals[(size_t)nb->nrem * S_COUNT + s] = nb->v[s][i];
        nb->nrem++;
        nb->cells[i] = nb->cells[last];
        for (int s = 0; s < S_COUNT; ++s) nb->v[s][i] = nb->v[s][last];
        nb->n = last;
        nb->fillers--;
        available -= single;
        removed++;
    }
    nb->total_filler_area = available;
    if (area - single * removed > single) {
        set_target_density(nb, td);
        nb->R.movable_area = (int64_t)((float)nb->R.white_space_area * nb->R.target_density);
    }
}

/* restoreRemovedFillers: each filler taken out since the last minimum back at the end, in the
 * order taken, with its values. */
static void restore_removed_fillers(orc_nb* nb) {
    const int64_t single = (int64_t)nb->R.filler_dx * nb->R.filler_dy;
    for (int k = 0; k < nb->nrem; ++k) {
        nb->cells[nb->n] = nb->rem_cell[k];
        for (int s = 0; s < S_COUNT; ++s) nb->v[s][nb->n] = nb->rem_vals[(size_t)k * S_COUNT + s];
        nb->n++;
        nb->fillers++;
        nb->total_filler_area += single;
    }
    nb->nrem = 0;
}

/* applyNetWeights: each net's weight from the map, at least 1; the count above 1. */
static int apply_net_weights(orc_np* np) {
    int w = 0;
    for (int e = 0; e < np->nnets; ++e) {
        np->nw[e] = np->rt.net_weight[e] > 1.0f ? np->rt.net_weight[e] : 1.0f;
        if (np->nw[e] > 1.0f) ++w;
    }
    np->rt.weighted = w;
    return w;
}

/* revertToMinCongestion. */
static void revert_to_min_congestion(orc_np* np) {
    if (np->rt.revert_count < 1) return;
    for (int g = 0; g < np->ncells; ++g) {
        if (!is_std(np, g)) continue;
        int32_t box[4] = {np->lx[g], np->ly[g], np->ux[g], np->uy[g]}, nbx[4];
        orc_set_size(box, np->rt.min_rc_dxy[2 * (size_t)g], np->rt.min_rc_dxy[2 * (size_t)g + 1], nbx);
        np->lx[g] = nbx[0];
        np->ly[g] = nbx[1];
        np->ux[g] = nbx[2];
        np->uy[g] = nbx[3];
        np->dxy[2 * (size_t)g] = nbx[2] - nbx[0];
        np->dxy[2 * (size_t)g + 1] = nbx[3] - nbx[1];
    }
    for (int b = 0; b < np->nregions; ++b) {
        orc_nb* nb = &np->nbs[b];
        set_target_density(nb, np->rt.min_rc_target_density[b]);
        restore_removed_fillers(nb);
        update_density_size_nb(np, nb);
        np->rt.accumulated[b] = np->rt.min_rc_inflated[b];
    }
    memcpy(np->rt.net_weight, np->rt.min_rc_net_weight, sizeof(float) * np->nnets);
    apply_net_weights(np);
    memset(np->rt.inflated_delta, 0, sizeof(int64_t) * np->nregions);
}

/* getTile: origin (0, 0), the last column and row clamped; -1 left of or below the grid. */
static int64_t tile_at(const orc_np* np, int32_t x, int32_t y) {
    int32_t ix = x / np->rt.tile, iy = y / np->rt.tile;
    if (ix > np->rt.cnt_x - 1) ix = np->rt.cnt_x - 1;
    if (iy > np->rt.cnt_y - 1) iy = np->rt.cnt_y - 1;
    if (ix < 0 || iy < 0) return -1;
    return (int64_t)iy * np->rt.cnt_x + ix;
}

/* RouteBase::measure's RUDY half (calculateRudyTiles + updateRudyAverage) at the placement now,
 * with each net's demand in the congested tiles (getNetDemandInTiles over the tiles of inflation
 * ratio above 1, route_base.rs:200-218, 381-400). */
static float measure_rudy(orc_np* np) {
    for (int e = 0; e < np->nnets; ++e) update_box(np, e);
    orc_rudy(np->rt.die, np->rt.cnt_x, np->rt.cnt_y, np->rt.tile, np->rt.wire_width, np->nnets, np->nbox, NULL, np->rt.demand, np->rt.blockage);
    int32_t nc = 0;
    for (int b = 0; b < np->nregions; ++b)
        for (int i = 0; i < np->nbs[b].n; ++i) nc += is_std(np, np->nbs[b].cells[i]);
    int32_t* centres = (int32_t*)xcalloc(2 * (size_t)(nc ? nc : 1), 4);
    nc = 0;
    for (int b = 0; b < np->nregions; ++b)
        for (int i = 0; i < np->nbs[b].n; ++i) {
            const int g = np->nbs[b].cells[i];
            if (!is_std(np, g)) continue;
            centres[2 * nc] = d_cx(np, g);
            centres[2 * nc + 1] = d_cy(np, g);
            nc++;
        }
    double st[8];
    orc_rudy_tiles(np->rt.cnt_x, np->rt.cnt_y, np->rt.tile, np->rt.demand, np->rt.blockage, nc, centres, np->rt.vars, np->rt.inflation_ratio,
                   np->rt.actionable, st);
    free(centres);
    np->rt.last_overflowed = (int32_t)st[1];
    np->rt.last_actionable = (int32_t)st[2];
    /* each net's demand in the congested tiles, in the visit's order (x outer, y inner) */
    const int32_t cx = np->rt.cnt_x, cy = np->rt.cnt_y, tile = np->rt.tile;
    for (int e = 0; e < np->nnets; ++e) {
        const int32_t* bx = np->nbox + 4 * (size_t)e;
        const int32_t xl = bx[0], yl = bx[1], xh = bx[2], yh = bx[3];
        float d = 0.0f;
        if (!(xl > xh || yl > yh) && (int64_t)(xh - xl) * (int64_t)(yh - yl) != 0) {
            const int64_t net_area = (int64_t)(xh - xl) * (int64_t)(yh - yl);
            const float cong = ((float)((xh - xl) + (yh - yl)) * (float)np->rt.wire_width) / (float)net_area;
            int32_t x0 = (xl - np->rt.die[0]) / tile, x1 = (xh - np->rt.die[0]) / tile, y0 = (yl - np->rt.die[1]) / tile, y1 = (yh - np->rt.die[1]) / tile;
            if (x0 < 0) x0 = 0;
            if (y0 < 0) y0 = 0;
            if (x1 > cx - 1) x1 = cx - 1;
            if (y1 > cy - 1) y1 = cy - 1;
            for (int32_t x = x0; x <= x1; ++x)
                for (int32_t y = y0; y <= y1; ++y) {
                    int32_t r[4];
                    orc_rudy_rect(np->rt.die, cx, cy, tile, x, y, r);
                    if (!(r[2] > xl && r[0] < xh && r[3] > yl && r[1] < yh)) continue;
                    if (!(np->rt.inflation_ratio[(size_t)y * cx + x] > 1.0f)) continue;
                    const int64_t inter = (int64_t)((xh < r[2] ? xh : r[2]) - (xl > r[0] ? xl : r[0])) *
                                          (int64_t)((yh < r[3] ? yh : r[3]) - (yl > r[1] ? yl : r[1]));
                    const int64_t tile_area = (int64_t)(r[2] - r[0]) * (int64_t)(r[3] - r[1]);
                    d += cong * ((float)inter / (float)tile_area) * 100.0f;
                }
        }
        np->rt.net_demand[e] = d;
    }
    return (float)st[7];
}

/* scaleInflationToBudget's trim (nesterov.rs:3950-3977): the pass may add a sixth of the budget or
 * what is left of it; when the cells ask for more (`wanted`, the sum of area x (ratio - 1)) every
 * inflated tile's excess over 1 is scaled by allowed / wanted (in double).  Returns whether the
 * budget is then spent.  Pinned by hand-computed cases (tests/test_oracle_route.py). */
int32_t orc_budget_trim(double budget, double spent, double wanted, int32_t ntiles, float* inflated_ratio) {
    const double allowed = budget / 6.0 < budget - spent ? budget / 6.0 : budget - spent;
    if (wanted <= 0.0 || wanted <= allowed) return spent + wanted >= budget;
    const double scale = allowed / wanted > 0.0 ? allowed / wanted : 0.0;
    for (int32_t t = 0; t < ntiles; ++t)
        if (inflated_ratio[t] > 1.0f) inflated_ratio[t] = (float)(1.0 + scale * ((double)inflated_ratio[t] - 1.0));
    return spent + allowed >= budget;
}

/* scaleInflationToBudget: the area this pass asks for against a sixth of the budget (or what is
 * left); over it every inflated tile's excess scaled down; whether the budget is spent. */
static int scale_inflation_to_budget(orc_np* np) {
    if (np->rt.original_movable_area <= 0) return 0;
    const double budget = (double)np->rt.original_movable_area * (double)np->o.routability_max_inflation_total;
    int64_t tot = 0;
    for (int b = 0; b < np->nregions; ++b) tot += np->rt.accumulated[b];
    const double spent = (double)tot;
    double wanted = 0.0;
    for (int b = 0; b < np->nregions; ++b)
        for (int i = 0; i < np->nbs[b].n; ++i) {
            const int g = np->nbs[b].cells[i];
            if (!is_std(np, g)) continue;
            const int64_t t = tile_at(np, d_cx(np, g), d_cy(np, g));
            if (t < 0) continue;
            const float ratio = np->rt.inflated_ratio[t];
            if ((double)ratio <= 1.0) continue;
            const double area = (double)(np->ux[g] - np->lx[g]) * (double)(np->uy[g] - np->ly[g]);
            wanted += area * ((double)ratio - 1.0);
        }
    return orc_budget_trim(budget, spent, wanted, np->rt.cnt_x * np->rt.cnt_y, np->rt.inflated_ratio);
}

/* weightCongestedNets' ranking (route_base.rs:381-426): -1 when no net was weighted. */
typedef struct {
    int32_t net;
    float d;
} net_dem;
static int by_demand(const void* a, const void* b) {
    const net_dem *x = (const net_dem*)a, *y = (const net_dem*)b;
    if (x->d != y->d) return x->d > y->d ? -1 : 1;
    return x->net < y->net ? -1 : (x->net > y->net ? 1 : 0);
}
/* weightCongestedNets' weights (route_base.rs:400-424) from each net's demand in the congested
 * tiles (nnets of them; 0: none): the nets with some, ranked by demand (largest first, ties in the
 * nets' order), the top `pct` % (at least one, ceil((float)n * pct) / 100) weighted from wmax at
 * the worst down to 1 at the cutoff (wmax throughout when they all tie); 0 for every other net.
 * Returns the count of nets with demand, -1 when none or when weighting is off.  Pinned by
 * hand-computed cases (tests/test_oracle_route.py). */
int32_t orc_congested_net_weights(int32_t nnets, const float* demand, float pct, float wmax, float* weight) {
    memset(weight, 0, sizeof(float) * (size_t)(nnets > 0 ? nnets : 0));
    if (wmax <= 1.0f || pct <= 0.0f || nnets == 0) return -1;
    net_dem* d = (net_dem*)xcalloc(nnets, sizeof(net_dem));
    size_t n = 0;
    for (int e = 0; e < nnets; ++e)
        if (demand[e] > 0.0f) {
            d[n].net = e;
            d[n].d = demand[e];
            n++;
        }
    if (n == 0) {
        free(d);
        return -1;
    }
    qsort(d, n, sizeof(net_dem), by_demand);
    const size_t share = (size_t)ceil((double)((float)n * pct) / 100.0);
    const size_t count = n < (share > 1 ? share : 1) ? n : (share > 1 ? share : 1);
    const float worst = d[0].d, cutoff = d[count - 1].d;
    for (size_t i = 0; i < count; ++i)
        weight[d[i].net] = worst > cutoff ? 1.0f + (wmax - 1.0f) * (d[i].d - cutoff) / (worst - cutoff) : wmax;
    free(d);
    return (int)n;
}

static int weight_congested_nets(orc_np* np) {
    return orc_congested_net_weights(np->nnets, np->rt.net_demand, np->o.routability_congested_nets_percentage,
                                     np->o.routability_net_weight_max, np->rt.net_weight);
}

static void push_pass(orc_np* np, gpl_route_pass* row, int need, int revert) {
    int64_t add = 0;
    for (int b = 0; b < np->nregions; ++b) add += np->rt.inflated_delta[b];
    row->inflation = add;
    row->target_density = np->nbs[0].R.target_density;
    row->fillers = np->nbs[0].fillers;
    row->weighted_nets = np->rt.weighted;
    row->need = need;
    row->revert = revert;
    if (np->rt.npasses == np->rt.cap_passes) {
        np->rt.cap_passes = np->rt.cap_passes ? 2 * np->rt.cap_passes : 16;
        np->rt.passes = (gpl_route_pass*)realloc(np->rt.passes, sizeof(gpl_route_pass) * np->rt.cap_passes);
    }
    np->rt.passes[np->rt.npasses++] = *row;
}

/* RouteBase::routability: *need (is_routability_need), *revert (is_revert_init_needed). */
static void routability(orc_np* np, int revert_count, int* need, int* revert) {
    np->rt.revert_count += 1;
    const float cur_rc = measure_rudy(np);
    gpl_route_pass row;
    memset(&row, 0, sizeof row);
    row.iter = np->iter_now;
    row.rc = cur_rc;
    row.overflowed_tiles = np->rt.last_overflowed;
    row.actionable_tiles = np->rt.last_actionable;
    if (cur_rc < np->o.routability_target_rc) { /* GPL-0050 */
        push_pass(np, &row, 0, 0);
        memset(np->rt.inflated_delta, 0, sizeof(int64_t) * np->nregions);
        *need = 0;
        *revert = 0;
        return;
    }
    if ((np->rt.min_rc - cur_rc) > np->rt.min_rc * 0.005f) { /* GPL-0048 */
        np->rt.min_rc = cur_rc;
        np->rt.min_rc_violated = 0;
        for (int b = 0; b < np->nregions; ++b) {
            np->rt.min_rc_target_density[b] = np->nbs[b].R.target_density;
            np->nbs[b].nrem = 0;
        }
        for (int g = 0; g < np->ncells; ++g)
            if (is_std(np, g)) {
                np->rt.min_rc_dxy[2 * (size_t)g] = np->ux[g] - np->lx[g];
                np->rt.min_rc_dxy[2 * (size_t)g + 1] = np->uy[g] - np->ly[g];
            }
        memcpy(np->rt.min_rc_inflated, np->rt.accumulated, sizeof(int64_t) * np->nregions);
        memcpy(np->rt.min_rc_net_weight, np->rt.net_weight, sizeof(float) * np->nnets);
    } else {
        np->rt.min_rc_violated += 1; /* GPL-0049 */
    }
    const size_t T = (size_t)np->rt.cnt_x * np->rt.cnt_y;
    for (size_t t = 0; t < T; ++t) np->rt.inflated_ratio[t] = np->rt.inflation_ratio[t] > 1.0f ? np->rt.inflation_ratio[t] : 1.0f;
    const int budget_spent = scale_inflation_to_budget(np);
    for (int b = 0; b < np->nregions; ++b) {
        orc_nb* nb = &np->nbs[b];
        np->rt.inflated_delta[b] = 0;
        for (int i = 0; i < nb->n; ++i) {
            const int g = nb->cells[i];
            if (!is_std(np, g)) continue;
            const int64_t t = tile_at(np, d_cx(np, g), d_cy(np, g));
            if (t < 0) continue;
            const float ratio = np->rt.inflated_ratio[t];
            if ((double)ratio <= 1.0) continue;
            const int64_t prev = (int64_t)(np->ux[g] - np->lx[g]) * (int64_t)(np->uy[g] - np->ly[g]);
            const float root = sqrtf(ratio);
            const int32_t nx = (int32_t)roundf((float)(np->ux[g] - np->lx[g]) * root), ny = (int32_t)roundf((float)(np->uy[g] - np->ly[g]) * root);
            int32_t box[4] = {np->lx[g], np->ly[g], np->ux[g], np->uy[g]}, nbx[4];
            orc_set_size(box, nx, ny, nbx);
            np->lx[g] = nbx[0];
            np->ly[g] = nbx[1];
            np->ux[g] = nbx[2];
            np->uy[g] = nbx[3];
            np->dxy[2 * (size_t)g] = nbx[2] - nbx[0];
            np->dxy[2 * (size_t)g + 1] = nbx[3] - nbx[1];
            np->rt.inflated_delta[b] += (int64_t)(nbx[2] - nbx[0]) * (int64_t)(nbx[3] - nbx[1]) - prev;
        }
        np->rt.accumulated[b] += np->rt.inflated_delta[b];
        cut_filler_cells(np, nb, np->rt.inflated_delta[b]);
        if (nb->R.target_density > np->o.routability_max_density) { /* GPL-0053 */
            revert_to_min_congestion(np);
            push_pass(np, &row, 0, 1);
            *need = 0;
            *revert = 1;
            return;
        }
    }
    if (revert_count >= np->o.routability_max_passes || np->rt.min_rc_violated >= 3) { /* GPL-0091, GPL-0054 */
        revert_to_min_congestion(np);
        push_pass(np, &row, 0, 1);
        *need = 0;
        *revert = 1;
        return;
    }
    for (int b = 0; b < np->nregions; ++b) {
        update_areas(np, &np->nbs[b]);
        update_density_size_nb(np, &np->nbs[b]);
    }
    weight_congested_nets(np);
    apply_net_weights(np);
    push_pass(np, &row, !budget_spent, 1);
    memset(np->rt.inflated_delta, 0, sizeof(int64_t) * np->nregions);
    if (budget_spent) revert_to_min_congestion(np); /* GPL-0097 */
    *need = !budget_spent;
    *revert = 1;
}

static int is_placement_settled(const orc_np* np) {
    for (int b = 0; b < np->nregions; ++b) {
        const orc_nb* nb = &np->nbs[b];
        if (!(nb->peak_coordi_distance > 0.0f && nb->coordi_distance <= 0.6f * nb->peak_coordi_distance)) return 0;
    }
    return 1;
}

/* tryRoutabilityDivergeRecovery: 1 when it reverted to the routability snapshot. */
static int try_routability_diverge_recovery(orc_np* np, float* cur_a) {
    if (!np->rt.saved || np->rt.attempts >= 2) return 0;
    if (!np->rt.need && (np->diverge_saved || np->diverge_reverts > 0)) return 0;
    np->rt.attempts = np->rt.need ? np->rt.attempts + 1 : 2;
    const int last = np->rt.attempts == 2;
    if (np->rt.need) revert_to_min_congestion(np);
    if (last) np->rt.need = 0;
    np->wl_coef_x = np->rt.coef_x;
    np->wl_coef_y = np->rt.coef_y;
    update_wirelength_force_wa(np);
    for (int b = 0; b < np->nregions; ++b) {
        revert_to_snapshot(np, &np->nbs[b]);
        np->nbs[b].min_sum_overflow = 1e30f;
        np->nbs[b].hpwl_with_min_sum_overflow = 1e30f;
    }
    *cur_a = 1.0f;
    np->rt.settle_start = -1;
    np->min_hpwl = INT64_MAX;
    np->is_min_hpwl = 0;
    return 1;
}

/* The loop's routability step after the divergence test: the routability snapshot, then a pass
 * once the overflow reached its end and the cells settled. */
static void routability_step(orc_np* np, int iter, float* cur_a) {
    if (!np->rt.saved && np->o.routability_snapshot_overflow >= np->average_overflow_unscaled) {
        np->rt.coef_x = np->wl_coef_x;
        np->rt.coef_y = np->wl_coef_y;
        np->rt.a = *cur_a;
        np->rt.saved = 1;
        for (int b = 0; b < np->nregions; ++b) save_snapshot(&np->nbs[b]);
    }
    const int active = np->rt.need && np->average_overflow_unscaled <= np->o.routability_end_overflow;
    const int settled = active && is_placement_settled(np);
    if (active && !settled && np->rt.settle_start == -1)
        np->rt.settle_start = iter;
    else if (settled && np->rt.settle_start != -1)
        np->rt.settle_start = -1;
    if (!settled) return;
    np->rt.reverts += 1;
    int need = 0, revert = 0;
    routability(np, np->rt.reverts, &need, &revert);
    np->rt.need = need;
    np->min_hpwl = INT64_MAX;
    np->is_min_hpwl = 0;
    if (need || revert) {
        *cur_a = np->rt.a;
        np->wl_coef_x = np->rt.coef_x;
        np->wl_coef_y = np->rt.coef_y;
        update_wirelength_force_wa(np);
        for (int b = 0; b < np->nregions; ++b) {
            revert_to_snapshot(np, &np->nbs[b]);
            np->nbs[b].min_sum_overflow = 1e30f;
            np->nbs[b].hpwl_with_min_sum_overflow = 1e30f;
        }
    }
}

/* The routability record: totals (passes, routability iterations, final and minimum congestion,
 * inflation, original area) into s, the pass rows into rows (up to cap); returns the pass count. */
int32_t orc_route_stats(orc_np* np, gpl_route_stats* s, gpl_route_pass* rows, int32_t cap) {
    if (s) {
        memset(s, 0, sizeof *s);
        s->passes = np->rt.npasses;
        s->route_iters = np->rt.route_iters;
        s->final_rc = np->rt.final_rc;
        s->min_rc = np->rt.min_rc;
        for (int b = 0; b < np->nregions && np->rt.accumulated; ++b) s->total_inflation += np->rt.accumulated[b];
        s->original_area = np->rt.original_movable_area;
    }
    for (int32_t i = 0; rows && i < np->rt.npasses && i < cap; ++i) rows[i] = np->rt.passes[i];
    return np->rt.npasses;
}

/* nesterov.rs:3107 checkConvergence (report lines aside). */
static int check_convergence(orc_np* np, int b) {
    orc_nb* nb = &np->nbs[b];
    if (nb->is_converged) return 1;
    if (nb->sum_overflow_unscaled > np->o.overflow) return 0;
    if (np->rt.on) np->rt.final_rc = measure_rudy(np); /* GPL-1005 */
    for (int k = 0; k < nb->n; ++k) {
        int g = nb->cells[k];
        if (np->flags[g] & GPL_CELL_INSTANCE) np->flags[g] |= GPL_CELL_LOCKED;
    }
    nb->is_converged = 1;
    return 1;
}

static void push_history(orc_np* np, int iter) {
    if (np->nhist == np->cap_hist) {
        np->cap_hist = np->cap_hist ? 2 * np->cap_hist : 1024;
        np->hist = (gpl_iter_row*)realloc(np->hist, sizeof(gpl_iter_row) * np->cap_hist);
    }
    gpl_iter_row* r = &np->hist[np->nhist++];
    r->iter = iter;
    r->trials = np->trials;
    r->overflow = np->average_overflow_unscaled;
    r->density_penalty = np->nbs[0].density_penalty;
    r->step_length = np->nbs[0].step_length;
    r->wl_coef = np->wl_coef_x;
    r->hpwl = np->nbs[0].prev_hpwl;
    for (int b = 0; b < np->nregions; ++b) {
        if (np->nrrows == np->cap_rrows) {
            np->cap_rrows = np->cap_rrows ? 2 * np->cap_rrows : 4096;
            np->rrows = (float*)realloc(np->rrows, sizeof(float) * 8 * (size_t)np->cap_rrows);
        }
        float* q = np->rrows + 8 * (size_t)np->nrrows++;
        q[0] = (float)iter;
        q[1] = (float)b;
        q[2] = np->nbs[b].sum_overflow_unscaled;
        q[3] = np->nbs[b].density_penalty;
        q[4] = np->nbs[b].step_length;
        q[5] = (float)np->nbs[b].is_converged;
        q[6] = np->nbs[b].wl_grad_sum;
        q[7] = np->nbs[b].density_grad_sum;
    }
}

/* Every region's state after each iteration (8 floats a row: iteration, region, overflow, penalty,
 * step length, converged, WA gradient sum, density gradient sum); returns the number of rows. */
int32_t orc_region_rows(orc_np* np, float* out, int32_t cap) {
    int32_t n = np->nrrows < cap ? np->nrrows : cap;
    if (out && n > 0) memcpy(out, np->rrows, sizeof(float) * 8 * (size_t)n);
    return np->nrrows;
}

/* nesterov.rs:4484-4486 the loop's momentum: a' = (1 + sqrt(4 a^2 + 1)) / 2 in double, and the
 * look-ahead coefficient (a - 1) / a'.  Pinned by the closed form of Nesterov's sequence
 * (a_k^2 - a_k = a_{k-1}^2, tests/test_oracle_control.py). */
float orc_momentum(float prev_a, float* coeff) {
    float a = (float)((1.0 + sqrt(4.0 * (double)prev_a * (double)prev_a + 1.0)) * 0.5);
    *coeff = (float)(((double)prev_a - 1.0) / (double)a);
    return a;
}

/* nesterov.rs:4440 resume_nesterov_place (no routability, timing or penalty guard).  Runs at most
 * `iter_budget` iterations (-1: no limit) and stops early once `time_budget_s` of wall time has
 * passed (<= 0: no limit) -- the bench's bounded CPU sample.  Returns the status; *iter_out is
 * the final iteration. */
int32_t orc_run(orc_np* np, int32_t start_iter, int32_t iter_budget, double time_budget_s, int32_t* iter_out) {
    double t0 = now_s();
    if (np->num_region_diverged > 0) return np->diverge_code;
    for (int b = 0; b < np->nregions; ++b) {
        orc_nb* nb = &np->nbs[b];
        nb->max_phi_coef_changed = 0;
        nb->min_sum_overflow = 1e30f;
        nb->hpwl_with_min_sum_overflow = 1e30f;
        nb->is_converged = 0;
    }
    np->diverge_saved = 0;
    np->diverge_reverts = 0;
    np->diverge_hpwl = 0;
    np->rt.saved = 0;
    np->rt.attempts = 0;
    np->rt.route_iters = 0;
    np->rt.reverts = 0;
    float cur_a = 1.0f;
    int iter = start_iter, ran = 0;
    np->converged = 0;
    while (iter < np->max_iter) {
        if (iter_budget >= 0 && ran >= iter_budget) break;
        if (time_budget_s > 0.0 && now_s() - t0 > time_budget_s) break;
        float coeff = 0.0f;
        cur_a = orc_momentum(cur_a, &coeff);
        np->iter_now = iter;
        do_back_tracking(np, coeff);
        np->total_trials += np->trials;
        for (int b = 0; b < np->nregions; ++b) nesterov_adjust_phi(&np->nbs[b]);
        if (np->num_region_diverged > 0) break;
        np_update_next_iter(np, iter);
        push_history(np, iter);
        ran++;
        np->iter_done = ran;
        /* is_routability_gpl_iter: each such iteration extends the limit by one */
        if (np->rt.on && np->rt.saved && np->average_overflow_unscaled > np->o.routability_end_overflow) {
            np->rt.route_iters += 1;
            np->max_iter += 1;
        }
        int act = np_is_diverged(np, &cur_a);
        if (act == 2) break;
        if (act == 1) {
            iter++;
            continue;
        }
        if (np->rt.on) routability_step(np, iter, &cur_a);
        int conv = 0;
        for (int b = 0; b < np->nregions; ++b) conv += check_convergence(np, b);
        if (conv == np->nregions) {
            np->converged = 1;
            np->t_converge = np->t_init + (now_s() - t0);
            break;
        }
        iter++;
    }
    if (orc_max_iter_takes_snapshot(iter, np->max_iter, np->o.disable_revert_if_diverge, np->diverge_saved, np->diverge_reverts,
                                    np->diverge_hpwl, np->diverge_snapshot_overflow, np->average_overflow_unscaled, orc_np_hpwl(np)))
        revert_to_diverge_snapshot(np);
    np->t_loop = now_s() - t0;
    if (iter_out) *iter_out = iter;
    if (np->num_region_diverged > 0) return np->diverge_code;
    return 0;
}

/* Loop statistics, in the C ABI's layout. */
void orc_stats(orc_np* np, gpl_stats* s) {
    memset(s, 0, sizeof *s);
    s->iter = np->nhist ? np->hist[np->nhist - 1].iter : -1;
    s->converged = np->converged;
    s->status = np->num_region_diverged > 0 ? np->diverge_code : 0;
    s->total_trials = np->total_trials;
    s->hpwl = orc_np_hpwl(np);
    s->overflow = np->average_overflow_unscaled;
    s->overflow_scaled = np->average_overflow;
    s->density_penalty = np->nbs[0].density_penalty;
    s->step_length = np->nbs[0].step_length;
    s->wl_coef = np->wl_coef_x;
    s->base_wl_coef = np->base_wl_coef;
    s->reference_hpwl = np->reference_hpwl;
    s->wl_grad_sum = np->nbs[0].wl_grad_sum;
    s->density_grad_sum = np->nbs[0].density_grad_sum;
    s->num_diverge_reverts = np->diverge_reverts;
    s->init_seconds = np->t_init;
    s->loop_seconds = np->t_loop;
    s->first_below_target_seconds = np->t_converge;
}

int32_t orc_history(orc_np* np, gpl_iter_row* rows, int32_t cap) {
    int32_t n = np->nhist < cap ? np->nhist : cap;
    memcpy(rows, np->hist, sizeof(gpl_iter_row) * n);
    return np->nhist;
}

/* Region 0's per-cell vectors by slot (the C ABI's GPL_SLOT_* numbering) and bins. */
void orc_get_slot(orc_np* np, int32_t b, int32_t slot, float* xy) {
    static const int map[] = {S_CUR_SLP, S_NEXT_SLP, S_PREV_SLP, S_CUR, S_NEXT};
    memcpy(xy, np->nbs[b].v[map[slot]], sizeof(fpt) * np->nbs[b].n);
}
void orc_get_grad(orc_np* np, int32_t b, int32_t which, float* xy) {
    /* 0 cur_slp_wl, 1 cur_slp_den, 2 cur_slp_sum, 3 prev_slp_sum, 4 next_slp_sum */
    static const int map[] = {S_CUR_SLP_WL, S_CUR_SLP_DEN, S_CUR_SLP_SUM, S_PREV_SLP_SUM, S_NEXT_SLP_SUM};
    memcpy(xy, np->nbs[b].v[map[which]], sizeof(fpt) * np->nbs[b].n);
}
void orc_get_bins(orc_np* np, int32_t b, float* density, float* field_x, float* field_y, int64_t* ovf2) {
    orc_nb* nb = &np->nbs[b];
    size_t M = (size_t)nb->R.cnt_x * nb->R.cnt_y;
    if (density) memcpy(density, nb->density, 4 * M);
    if (field_x) memcpy(field_x, nb->field_x, 4 * M);
    if (field_y) memcpy(field_y, nb->field_y, 4 * M);
    if (ovf2) {
        ovf2[0] = nb->overflow;
        ovf2[1] = nb->overflow_unscaled;
    }
}
void orc_get_scalars(orc_np* np, int32_t b, float* out) {
    orc_nb* nb = &np->nbs[b];
    out[0] = nb->step_length;
    out[1] = nb->density_penalty;
    out[2] = nb->wl_grad_sum;
    out[3] = nb->density_grad_sum;
    out[4] = nb->sum_overflow;
    out[5] = nb->sum_overflow_unscaled;
    out[6] = np->wl_coef_x;
    out[7] = np->base_wl_coef;
    out[8] = np->reference_hpwl;
    out[9] = nb->coordi_distance;
}
void orc_set_inject(orc_np* np, int32_t after_iter) { np->inject_nan_after = after_iter; }

/* The forced loop-control cases (see orc_np): kind 5 the zero WA sum at iteration `at`, kind 6 a
 * divergence at iteration `at` (a second call sets a second iteration); kind 0 clears both. */
void orc_set_inject_control(orc_np* np, int32_t kind, int32_t at) {
    if (kind == 0) {
        np->inject_zero_wl_at = np->inject_diverge_at = np->inject_diverge_at2 = -1;
    } else if (kind == 5) {
        np->inject_zero_wl_at = at;
    } else if (kind == 6) {
        if (np->inject_diverge_at < 0) np->inject_diverge_at = at;
        else np->inject_diverge_at2 = at;
    }
}

/* ---- per-kernel oracles on a caller-given state --------------------------------------------- */

/* The scatter of a region's cells at coordinates xy (density centres truncated, as
 * update_gcell_density_center_location does), with the integer bin areas, the float density and
 * the two overflows: the reference's f32-through-i64 sums (ovf_i64[2]) and exact FP64 sums of
 * the same per-bin terms (ovf_f64[2]).  The cells' density sizes come from the problem (after
 * the first centring the boxes are 2*(d/2) wide). */
void orc_scatter(orc_np* np, int32_t b, const float* xy, int64_t* inst, int64_t* filler, float* density, int64_t* ovf_i64, double* ovf_f64) {
    orc_nb* nb = &np->nbs[b];
    update_center_location(np, nb, (const fpt*)xy);
    size_t M = (size_t)nb->R.cnt_x * nb->R.cnt_y;
    if (inst) memcpy(inst, nb->inst, 8 * M);
    if (filler) memcpy(filler, nb->filler, 8 * M);
    if (density) memcpy(density, nb->density, 4 * M);
    if (ovf_i64) {
        ovf_i64[0] = nb->overflow;
        ovf_i64[1] = nb->overflow_unscaled;
    }
    if (ovf_f64) {
        double o = 0.0, ou = 0.0;
        float td = nb->R.target_density;
        for (int y = 0; y < nb->R.cnt_y; ++y)
            for (int x = 0; x < nb->R.cnt_x; ++x) {
                size_t k = (size_t)y * nb->R.cnt_x + x;
                int64_t area = (int64_t)(nb->bin_x[x + 1] - nb->bin_x[x]) * (int64_t)(nb->bin_y[y + 1] - nb->bin_y[y]);
                float scaled = (float)area * td;
                float t = (float)nb->inst[k] + (float)nb->non_place[k] - scaled;
                float tu = (float)nb->inst[k] + (float)nb->non_place_unscaled[k] - scaled;
                o += t > 0.0f ? (double)t : 0.0;
                ou += tu > 0.0f ? (double)tu : 0.0;
            }
        ovf_f64[0] = o;
        ovf_f64[1] = ou;
    }
}

/* Sets the pins of a region's cells from density centres (ints), without touching the bins. */
void orc_place_pins(orc_np* np, int32_t b, const float* xy) {
    orc_nb* nb = &np->nbs[b];
    for (int i = 0; i < nb->n; ++i) set_density_center(np, nb->cells[i], (int)xy[2 * i], (int)xy[2 * i + 1]);
}

/* WA forward at the current pins: faithful float net sums and boxes, FP64 net sums, HPWL. */
void orc_wa_forward(orc_np* np, float coef_x, float coef_y, float* sums, int32_t* box, double* sums_f64, int64_t* hpwl) {
    np->wl_coef_x = coef_x;
    np->wl_coef_y = coef_y;
    update_wirelength_force_wa(np);
    if (sums) memcpy(sums, np->nsum, sizeof(float) * 8 * (size_t)np->nnets);
    if (box) memcpy(box, np->nbox, sizeof(int32_t) * 4 * (size_t)np->nnets);
    if (sums_f64) {
        for (int n = 0; n < np->nnets; ++n) {
            double s[8] = {0};
            int nlx = np->nbox[4 * n], nly = np->nbox[4 * n + 1], nux = np->nbox[4 * n + 2], nuy = np->nbox[4 * n + 3];
            for (int k = np->net_start[n]; k < np->net_start[n + 1]; ++k) {
                int p = np->net_pins[k], px = np->pcx[p], py = np->pcy[p];
                float e[4] = {(float)(nlx - px) * coef_x, (float)(px - nux) * coef_x, (float)(nly - py) * coef_y, (float)(py - nuy) * coef_y};
                for (int q = 0; q < 4; ++q)
                    if (e[q] > -300.0f) {
                        double v = (double)orc_fast_exp(e[q]);
                        s[2 * q] += v;
                        s[2 * q + 1] += (double)(q < 2 ? px : py) * v;
                    }
            }
            for (int q = 0; q < 8; ++q) sums_f64[8 * n + q] = s[q];
        }
    }
    if (hpwl) {
        int64_t t = 0;
        for (int n = 0; n < np->nnets; ++n) {
            int32_t* bb = &np->nbox[4 * n];
            if (bb[2] >= bb[0]) t += ((int64_t)bb[2] - bb[0]) + ((int64_t)bb[3] - bb[1]);
        }
        *hpwl = t;
    }
}

/* WA backward of a region's cells after orc_wa_forward: faithful float, and FP64 from the FP64
 * net sums (the same quotient and per-pin exponentials, accumulated in double). */
void orc_wa_backward(orc_np* np, int32_t b, float* wl, const double* sums_f64, double* wl_f64) {
    orc_nb* nb = &np->nbs[b];
    for (int i = 0; i < nb->n; ++i) {
        int g = nb->cells[i];
        fpt v = wl_gradient(np, g);
        if (wl) {
            wl[2 * i] = v.x;
            wl[2 * i + 1] = v.y;
        }
        if (wl_f64 && sums_f64) {
            double sx = 0.0, sy = 0.0;
            for (int k = np->pin_start[g]; k < np->pin_start[g + 1]; ++k) {
                int p = np->pins[k], n = np->pnet[p];
                const double* G = &sums_f64[8 * n];
                const float* q = &np->pexp[4 * p];
                double cx = np->wl_coef_x, cy = np->wl_coef_y, px = np->pcx[p], py = np->pcy[p];
                double g0 = 0, g1 = 0, g2 = 0, g3 = 0;
                if (np->has[p] & 1) g0 = (G[0] * q[0] * (1.0 - cx * px) + cx * q[0] * G[1]) / (G[0] * G[0]);
                if (np->has[p] & 2) g1 = (G[2] * q[1] * (1.0 + cx * px) - cx * q[1] * G[3]) / (G[2] * G[2]);
                if (np->has[p] & 4) g2 = (G[4] * q[2] * (1.0 - cy * py) + cy * q[2] * G[5]) / (G[4] * G[4]);
                if (np->has[p] & 8) g3 = (G[6] * q[3] * (1.0 + cy * py) - cy * q[3] * G[7]) / (G[6] * G[6]);
                sx += (g0 - g1) * np->nw[n];
                sy += (g2 - g3) * np->nw[n];
            }
            wl_f64[2 * i] = sx;
            wl_f64[2 * i + 1] = sy;
        }
    }
}

/* Density gather of a region's cells at their current density boxes over the given fields
 * (y-major): faithful float, and FP64 over FP64 fields. */
void orc_gather(orc_np* np, int32_t b, const float* field_x, const float* field_y, float* den, const double* fx64, const double* fy64, double* den_f64) {
    orc_nb* nb = &np->nbs[b];
    size_t M = (size_t)nb->R.cnt_x * nb->R.cnt_y;
    if (field_x) memcpy(nb->field_x, field_x, 4 * M);
    if (field_y) memcpy(nb->field_y, field_y, 4 * M);
    for (int i = 0; i < nb->n; ++i) {
        int g = nb->cells[i];
        if (den) {
            fpt f = density_gradient(np, nb, g);
            den[2 * i] = f.x;
            den[2 * i + 1] = f.y;
        }
        if (den_f64 && fx64 && fy64) {
            int x0, x1, y0, y1;
            didx_x(nb, np->dlx[g], np->dux[g], &x0, &x1);
            didx_y(nb, np->dly[g], np->duy[g], &y0, &y1);
            double sx = 0.0, sy = 0.0;
            for (int x = x0; x < x1; ++x)
                for (int y = y0; y < y1; ++y) {
                    size_t k = (size_t)y * nb->R.cnt_x + x;
                    double a = (double)overlap_density_area(nb->bin_x[x], nb->bin_y[y], nb->bin_x[x + 1], nb->bin_y[y + 1], np->dlx[g], np->dly[g], np->dux[g], np->duy[g]) * np->dscale[g];
                    sx += a * fx64[k];
                    sy += a * fy64[k];
                }
            den_f64[2 * i] = sx;
            den_f64[2 * i + 1] = sy;
        }
    }
}

/* The current density boxes and pin positions (for comparing state). */
void orc_get_dbox(orc_np* np, int32_t* dbox) {
    for (int g = 0; g < np->ncells; ++g) {
        dbox[4 * g] = np->dlx[g];
        dbox[4 * g + 1] = np->dly[g];
        dbox[4 * g + 2] = np->dux[g];
        dbox[4 * g + 3] = np->duy[g];
    }
}
void orc_get_pins(orc_np* np, int32_t* xy) {
    for (int p = 0; p < np->npins; ++p) {
        xy[2 * p] = np->pcx[p];
        xy[2 * p + 1] = np->pcy[p];
    }
}

/* ---- one net's WA model along one axis, for pinning the gradient expression ----------------- */

/* The WA wirelength W = X_max - X_min of one net's pin coordinates xs[0..n) (X_max / X_min the
 * exponentially weighted averages with shifts at the box, nesterov.rs:2354-2384) and per pin the
 * reference's gradient expression gmin - gmax (nesterov.rs:2392-2413) in double, with the
 * reference's fast exponential (use_fast) or the C library's exp.  With exp, gmin - gmax is
 * exactly -dW/dx_i, which the tests check against central differences; the force bar is ignored
 * here (every term kept). */
double orc_wa_net_f64(int32_t n, const double* xs, double coef, int32_t use_fast, double* grad) {
    double lx = xs[0], ux = xs[0];
    for (int i = 1; i < n; ++i) {
        if (xs[i] < lx) lx = xs[i];
        if (xs[i] > ux) ux = xs[i];
    }
    double s0n = 0, s1n = 0, s0x = 0, s1x = 0;
    double* en = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    double* ex = en + n;
    for (int i = 0; i < n; ++i) {
        double a = (lx - xs[i]) * coef, b = (xs[i] - ux) * coef;
        en[i] = use_fast ? (double)orc_fast_exp((float)a) : exp(a);
        ex[i] = use_fast ? (double)orc_fast_exp((float)b) : exp(b);
        s0n += en[i];
        s1n += xs[i] * en[i];
        s0x += ex[i];
        s1x += xs[i] * ex[i];
    }
    for (int i = 0; i < n && grad; ++i) {
        double gmin = (s0n * (en[i] * (1.0 - coef * xs[i])) + coef * en[i] * s1n) / (s0n * s0n);
        double gmax = (s0x * (ex[i] * (1.0 + coef * xs[i])) - coef * ex[i] * s1x) / (s0x * s0x);
        grad[i] = gmin - gmax;
    }
    free(en);
    return s1x / s0x - s1n / s0n;
}

/* ---- the differentiable objective (no reference counterpart), in FP64 ----------------------- */
/*
 * f(x) = sum_e w_e (WA_e,x + WA_e,y) + lambda * 1/2 sum_b rho_b phi_b over region 0's cells at the
 * continuous centres xy (problem order): WA with the exact exponential (the weighted-average
 * wirelength of nesterov.rs:2325-2388 without fast_exp), pins at centre + offset (pins of other
 * cells at their initial positions, ports fixed); rho_b = (sum_i overlap_b(i) scale_i [target if
 * macro] + non_place_b) / (area_b target) with each cell's density box of the problem's size
 * centred at its point (not truncated to integers); phi = K rho, K the Poisson solve.
 */
typedef struct {
    const gpl_problem* p;
    const gpl_region* R;
    const float* xy;
    double c;
} obj_ctx;

static void obj_pin(const obj_ctx* o, int32_t q, int32_t moved, int axis, double delta, double* px, double* py) {
    int32_t g = o->p->pin_cell[q];
    if (g >= o->R->cell_begin && g < o->R->cell_end) {
        int32_t k = g - o->R->cell_begin;
        *px = (double)o->xy[2 * k] + (double)o->p->pin_offset[2 * q];
        *py = (double)o->xy[2 * k + 1] + (double)o->p->pin_offset[2 * q + 1];
        if (g == moved) {
            if (axis == 0) *px += delta;
            else *py += delta;
        }
    } else {
        *px = (double)o->p->pin_xy[2 * q];
        *py = (double)o->p->pin_xy[2 * q + 1];
    }
}

/* one net's WA (x plus y) in double, with cell `moved` shifted by delta along axis */
static double obj_net_wa(const obj_ctx* o, int32_t e, int32_t moved, int axis, double delta) {
    const gpl_problem* p = o->p;
    int32_t s = p->net_pin_start[e], t = p->net_pin_start[e + 1];
    if (t - s < 1) return 0.0;
    double lx = DBL_MAX, ly = DBL_MAX, ux = -DBL_MAX, uy = -DBL_MAX;
    for (int32_t k = s; k < t; ++k) {
        double px, py;
        obj_pin(o, p->net_pins[k], moved, axis, delta, &px, &py);
        if (px < lx) lx = px;
        if (px > ux) ux = px;
        if (py < ly) ly = py;
        if (py > uy) uy = py;
    }
    double a0 = 0, a1 = 0, b0 = 0, b1 = 0, c0 = 0, c1 = 0, d0 = 0, d1 = 0;
    for (int32_t k = s; k < t; ++k) {
        double px, py;
        obj_pin(o, p->net_pins[k], moved, axis, delta, &px, &py);
        double ea = exp(o->c * (px - ux)), eb = exp(o->c * (lx - px));
        double ec = exp(o->c * (py - uy)), ed = exp(o->c * (ly - py));
        a0 += ea;
        a1 += px * ea;
        b0 += eb;
        b1 += px * eb;
        c0 += ec;
        c1 += py * ec;
        d0 += ed;
        d1 += py * ed;
    }
    return (a1 / a0 - b1 / b0) + (c1 / c0 - d1 / d0);
}

/* cell k's (region-local) density box at xy (+ delta along axis): lx, ly, ux, uy */
static void obj_box(const obj_ctx* o, int32_t k, int axis, double delta, double* b) {
    int32_t g = o->R->cell_begin + k;
    const int32_t* d = &o->p->cell_dbox[4 * (size_t)g];
    double w = (double)(d[2] - d[0]), h = (double)(d[3] - d[1]);
    double x = (double)o->xy[2 * k] + (axis == 0 ? delta : 0.0), y = (double)o->xy[2 * k + 1] + (axis == 1 ? delta : 0.0);
    b[0] = x - 0.5 * w;
    b[1] = y - 0.5 * h;
    b[2] = x + 0.5 * w;
    b[3] = y + 0.5 * h;
}

static double obj_cell_weight(const obj_ctx* o, int32_t k) {
    int32_t g = o->R->cell_begin + k;
    double w = (double)o->p->density_scale[g];
    uint8_t f = o->p->cell_flags[g];
    if ((f & GPL_CELL_INSTANCE) && (f & GPL_CELL_MACRO)) w *= (double)o->R->target_density;
    return w;
}

/* adds s * overlap(b, bin) / (area * target) into rho for every bin the box touches */
static void obj_add_box(const obj_ctx* o, const double* b, double s, double* rho) {
    const gpl_region* R = o->R;
    int x0 = (int)floor((b[0] - R->lx) / R->size_x) - 1, x1 = (int)floor((b[2] - R->lx) / R->size_x) + 1;
    int y0 = (int)floor((b[1] - R->ly) / R->size_y) - 1, y1 = (int)floor((b[3] - R->ly) / R->size_y) + 1;
    if (x0 < 0) x0 = 0;
    if (y0 < 0) y0 = 0;
    if (x1 > R->cnt_x - 1) x1 = R->cnt_x - 1;
    if (y1 > R->cnt_y - 1) y1 = R->cnt_y - 1;
    for (int y = y0; y <= y1; ++y) {
        double by0 = R->bin_y[y], by1 = R->bin_y[y + 1];
        double oy = fmin(b[3], by1) - fmax(b[1], by0);
        if (oy <= 0) continue;
        for (int x = x0; x <= x1; ++x) {
            double bx0 = R->bin_x[x], bx1 = R->bin_x[x + 1];
            double ox = fmin(b[2], bx1) - fmax(b[0], bx0);
            if (ox <= 0) continue;
            rho[(size_t)y * R->cnt_x + x] += s * ox * oy / ((bx1 - bx0) * (by1 - by0) * (double)R->target_density);
        }
    }
}

static double* obj_rho(const obj_ctx* o) {
    const gpl_region* R = o->R;
    size_t M = (size_t)R->cnt_x * R->cnt_y;
    double* rho = (double*)calloc(M, sizeof(double));
    for (int y = 0; y < R->cnt_y; ++y)
        for (int x = 0; x < R->cnt_x; ++x) {
            double a = (double)(R->bin_x[x + 1] - R->bin_x[x]) * (double)(R->bin_y[y + 1] - R->bin_y[y]);
            rho[(size_t)y * R->cnt_x + x] = (double)R->non_place[(size_t)y * R->cnt_x + x] / (a * (double)R->target_density);
        }
    for (int32_t k = 0; k < R->cell_end - R->cell_begin; ++k) {
        double b[4];
        obj_box(o, k, 0, 0.0, b);
        obj_add_box(o, b, obj_cell_weight(o, k), rho);
    }
    return rho;
}

/* The objective's value in FP64: *wa = sum_e w_e WA_e, *den = 1/2 sum rho phi. */
void orc_objective_value(const gpl_problem* p, const float* xy, float wl_coef, double* wa, double* den) {
    obj_ctx o = {p, &p->regions[0], xy, (double)wl_coef};
    double w = 0.0;
    for (int32_t e = 0; e < p->num_nets; ++e) w += (double)p->net_weight[e] * obj_net_wa(&o, e, -1, 0, 0.0);
    const gpl_region* R = o.R;
    size_t M = (size_t)R->cnt_x * R->cnt_y;
    double* rho = obj_rho(&o);
    double* phi = (double*)malloc(sizeof(double) * M);
    orc_fft_solve_f64d(R->cnt_x, R->cnt_y, R->wx, R->wy, rho, phi, NULL, NULL);
    double d = 0.0;
    for (size_t k = 0; k < M; ++k) d += rho[k] * phi[k];
    free(rho);
    free(phi);
    *wa = w;
    *den = 0.5 * d;
}

/*
 * Central differences in FP64 of the objective for the sampled cells (region-local indices):
 * out[4 i] = (f(x + h e_x) - f(x - h e_x)) / 2h, out[4 i + 1] the same along y, out[4 i + 2],
 * out[4 i + 3] = 1 where a box edge lies within h of a bin edge along that axis (the density term
 * has a kink there; the difference is not a derivative).  The wirelength part re-evaluates the
 * nets on the cell's pins; the density part is phi^T (rho(x+h) - rho(x-h)), exact for the
 * quadratic 1/2 rho^T K rho with K symmetric wherever rho is linear on [x-h, x+h] (no kink).
 */
void orc_objective_fd(const gpl_problem* p, const float* xy, float wl_coef, float lambda, int32_t ns, const int32_t* sample, double h,
                      double* out) {
    obj_ctx o = {p, &p->regions[0], xy, (double)wl_coef};
    const gpl_region* R = o.R;
    size_t M = (size_t)R->cnt_x * R->cnt_y;
    double* rho = obj_rho(&o);
    double* phi = (double*)malloc(sizeof(double) * M);
    orc_fft_solve_f64d(R->cnt_x, R->cnt_y, R->wx, R->wy, rho, phi, NULL, NULL);
    double* dr = (double*)calloc(M, sizeof(double));
    for (int32_t i = 0; i < ns; ++i) {
        int32_t k = sample[i], g = R->cell_begin + k;
        for (int axis = 0; axis < 2; ++axis) {
            double dw = 0.0;
            for (int32_t q = p->cell_pin_start[g]; q < p->cell_pin_start[g + 1]; ++q) {
                int32_t e = p->pin_net[p->cell_pins[q]];
                /* a net with several pins on this cell is visited once per pin: count it once */
                int32_t first = 1;
                for (int32_t q2 = p->cell_pin_start[g]; q2 < q; ++q2)
                    if (p->pin_net[p->cell_pins[q2]] == e) first = 0;
                if (!first) continue;
                dw += (double)p->net_weight[e] * (obj_net_wa(&o, e, g, axis, h) - obj_net_wa(&o, e, g, axis, -h));
            }
            double bp[4], bm[4], s = obj_cell_weight(&o, k);
            obj_box(&o, k, axis, h, bp);
            obj_box(&o, k, axis, -h, bm);
            obj_add_box(&o, bp, s, dr);
            obj_add_box(&o, bm, -s, dr);
            double dd = 0.0;
            int x0 = (int)floor((fmin(bp[0], bm[0]) - R->lx) / R->size_x) - 1, x1 = (int)floor((fmax(bp[2], bm[2]) - R->lx) / R->size_x) + 1;
            int y0 = (int)floor((fmin(bp[1], bm[1]) - R->ly) / R->size_y) - 1, y1 = (int)floor((fmax(bp[3], bm[3]) - R->ly) / R->size_y) + 1;
            if (x0 < 0) x0 = 0;
            if (y0 < 0) y0 = 0;
            if (x1 > R->cnt_x - 1) x1 = R->cnt_x - 1;
            if (y1 > R->cnt_y - 1) y1 = R->cnt_y - 1;
            for (int y = y0; y <= y1; ++y)
                for (int x = x0; x <= x1; ++x) {
                    size_t b = (size_t)y * R->cnt_x + x;
                    dd += phi[b] * dr[b];
                    dr[b] = 0.0;
                }
            out[4 * (size_t)i + axis] = (dw + (double)lambda * dd) / (2.0 * h);
            /* kink: an edge of the box along this axis within h of a bin edge (or the grid's) */
            const int32_t* edges = axis == 0 ? R->bin_x : R->bin_y;
            int32_t ne = (axis == 0 ? R->cnt_x : R->cnt_y) + 1;
            double lo = axis == 0 ? 0.5 * (bp[0] + bm[0]) : 0.5 * (bp[1] + bm[1]);
            double hi = axis == 0 ? 0.5 * (bp[2] + bm[2]) : 0.5 * (bp[3] + bm[3]);
            int kink = 0;
            for (int32_t j = 0; j < ne && !kink; ++j) {
                double ev = (double)edges[j];
                if (fabs(ev - lo) <= h || fabs(ev - hi) <= h) kink = 1;
            }
            out[4 * (size_t)i + 2 + axis] = kink;
        }
    }
    free(dr);
    free(rho);
    free(phi);
}

/* The Poisson solve as an operator on double densities (for the symmetry check). */
void orc_solve_phi_f64(int32_t nx, int32_t ny, const float* wx, const float* wy, const double* density, double* phi) {
    orc_fft_solve_f64d(nx, ny, wx, wy, density, phi, NULL, NULL);
}
