/*
 * TEST INFRASTRUCTURE -- part of the oracle.  Only tests/, the smoke check and the bench's CPU
 * baseline may call it; the product never does.
 *
 * Routability's congestion measurement restated from the reference (route_base.rs): the RUDY
 * grid and its wire width (Rudy::new / makeGrid, 143-174), the tiles' blockage from the router's
 * capacity reductions (getResourceReductions, 183-188), every net's demand in the tiles its
 * terminal box strictly overlaps, summed in float in the nets' order (calculate /
 * visitNetTiles, 191-253), the tile congestion (getTileCongestion, 256-264), the tile grid's
 * actionable tiles and inflation ratios (calculateRudyTiles, 443-488) and the top 0.5/1/2/5 %
 * averages with the weighted routing congestion (updateRudyAverage / top_averages, 493-536,
 * 661-675); and the per-cell half of a routability pass (nesterov.rs:3945-3960: the cells in
 * inflated tiles grown by the root of the ratio, set_size 3754-3760) with updateDensitySize
 * (2022-2036).  Single thread, the reference's order and float/double types throughout.
 *
 * Pinned by tests/test_oracle_route.py: the reference's own unit-test vectors (route_base.rs
 * 682-757: the wire width, the stretched last tile, strict overlap, blockage and congestion, the
 * top shares; nesterov.rs:4841-4845: set_size), the total-demand identity on random nets, and
 * hand-computed inflation, threshold, cap and density-size cases.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Rudy::new: 1.25 / sum(1 / pitch), truncated; 100 with no pitch. */
int32_t orc_rudy_wire_width(const int32_t* pitches, int32_t n) {
    double terms = 0.0;
    for (int32_t i = 0; i < n; ++i) terms += 1.0 / (double)pitches[i];
    return terms != 0.0 ? (int32_t)((1.0 / terms) * 1.25) : 100;
}

/* makeGrid: tile (x, y) (x-major index x * cnt_y + y) from the die's lower left; the last column
 * and row stretched by what the grid leaves of the die, measured from the origin. */
void orc_rudy_rect(const int32_t die[4], int32_t cnt_x, int32_t cnt_y, int32_t tile, int32_t x, int32_t y, int32_t* r) {
    const int32_t x_extra = die[2] - cnt_x * tile, y_extra = die[3] - cnt_y * tile;
    const int32_t cx = die[0] + x * tile, cy = die[1] + y * tile;
    r[0] = cx;
    r[1] = cy;
    r[2] = cx + tile + (x == cnt_x - 1 ? x_extra : 0);
    r[3] = cy + tile + (y == cnt_y - 1 ? y_extra : 0);
}

/* calculateRudy: blockage from (capacity, reduction) pairs (x-major; NULL: none), then every net's
 * demand (boxes xl, yl, xh, yh in the nets' order) into the tiles, float sums in that order. */
void orc_rudy(const int32_t die[4], int32_t cnt_x, int32_t cnt_y, int32_t tile, int32_t wire_width, int32_t nnets, const int32_t* boxes,
              const uint8_t* cap_red, float* demand, float* blockage) {
    const size_t n = (size_t)cnt_x * (size_t)cnt_y;
    for (size_t i = 0; i < n; ++i) {
        demand[i] = 0.0f;
        blockage[i] = 0.0f;
        if (cap_red) {
            const uint8_t cap = cap_red[2 * i], red = cap_red[2 * i + 1];
            if (cap == 0) {
                blockage[i] = 1.0f;
            } else {
                const float b = (float)red / (float)cap;
                blockage[i] = b < 1.0f ? b : 1.0f;
            }
        }
    }
    for (int32_t e = 0; e < nnets; ++e) {
        const int32_t xl = boxes[4 * e], yl = boxes[4 * e + 1], xh = boxes[4 * e + 2], yh = boxes[4 * e + 3];
        if (xl > xh || yl > yh) continue;
        const int64_t net_area = (int64_t)(xh - xl) * (int64_t)(yh - yl);
        if (net_area == 0) continue;
        const float hpwl = (float)((xh - xl) + (yh - yl));
        const float cong = (hpwl * (float)wire_width) / (float)net_area;
        int32_t min_x = (xl - die[0]) / tile, max_x = (xh - die[0]) / tile;
        int32_t min_y = (yl - die[1]) / tile, max_y = (yh - die[1]) / tile;
        if (min_x < 0) min_x = 0;
        if (min_y < 0) min_y = 0;
        if (max_x > cnt_x - 1) max_x = cnt_x - 1;
        if (max_y > cnt_y - 1) max_y = cnt_y - 1;
        for (int32_t x = min_x; x <= max_x; ++x)
            for (int32_t y = min_y; y <= max_y; ++y) {
                int32_t r[4];
                orc_rudy_rect(die, cnt_x, cnt_y, tile, x, y, r);
                if (!(r[2] > xl && r[0] < xh && r[3] > yl && r[1] < yh)) continue;
                const int32_t ixh = xh < r[2] ? xh : r[2], ixl = xl > r[0] ? xl : r[0];
                const int32_t iyh = yh < r[3] ? yh : r[3], iyl = yl > r[1] ? yl : r[1];
                const int64_t inter = (int64_t)(ixh - ixl) * (int64_t)(iyh - iyl);
                const int64_t tile_area = (int64_t)(r[2] - r[0]) * (int64_t)(r[3] - r[1]);
                const float ratio = (float)inter / (float)tile_area;
                demand[(size_t)x * cnt_y + y] += cong * ratio * 100.0f;
            }
    }
}

/* getTileCongestion (gpl's): demand over what the blockage leaves, infinite when nothing is. */
float orc_rudy_congestion(float demand, float blockage) {
    const float available = 1.0f - blockage;
    if (available <= 0.0f) return INFINITY;
    return (demand / 100.0f) / available;
}

static int desc(const void* a, const void* b) {
    const double x = *(const double*)a, y = *(const double*)b;
    return x < y ? 1 : (x > y ? -1 : 0);
}

/* top_averages: element i of the descending list joins a share when i < share * n; each sum over
 * ceil(share * n). */
void orc_top_averages(int32_t n, const double* desc_list, double* out4) {
    const double shares[4] = {0.005, 0.01, 0.02, 0.05};
    double sums[4] = {0.0, 0.0, 0.0, 0.0};
    const double nd = (double)n;
    for (int32_t i = 0; i < n; ++i)
        for (int s = 0; s < 4; ++s)
            if ((double)i < shares[s] * nd) sums[s] += desc_list[i];
    out4[0] = sums[0] / ceil(0.005 * nd);
    out4[1] = sums[1] / ceil(0.010 * nd);
    out4[2] = sums[2] / ceil(0.020 * nd);
    out4[3] = sums[3] / ceil(0.050 * nd);
}

/* calculateRudyTiles' tile half and updateRudyAverage.  The tile grid is y-major from the origin
 * (0, 0) at the RUDY tile size; a tile holding a standard cell's density centre (centres: x, y
 * pairs) is actionable unless its blockage reaches ignore_edge_ratio; an actionable tile at
 * congestion >= min_cong (> 0) takes min((c / min_cong)^coef, max_ratio), every other 1.
 * vars: min_cong, coef, max_ratio, ignore_edge_ratio, rc_k[4].  stats: total overflow, overflowed
 * tiles, actionable tiles counted, the four top averages, the weighted congestion (as float). */
void orc_rudy_tiles(int32_t cnt_x, int32_t cnt_y, int32_t tile, const float* demand, const float* blockage, int32_t ncentres,
                    const int32_t* centres, const float* vars, float* inflation_ratio, uint8_t* actionable, double* stats) {
    const size_t n = (size_t)cnt_x * (size_t)cnt_y;
    for (size_t t = 0; t < n; ++t) {
        actionable[t] = 0;
        inflation_ratio[t] = 1.0f;
    }
    for (int32_t i = 0; i < ncentres; ++i) {
        int32_t ix = centres[2 * i] / tile, iy = centres[2 * i + 1] / tile;
        if (ix > cnt_x - 1) ix = cnt_x - 1;
        if (iy > cnt_y - 1) iy = cnt_y - 1;
        if (ix < 0 || iy < 0) continue;
        actionable[(size_t)iy * cnt_x + ix] = 1;
    }
    const float min_cong = vars[0], coef = vars[1], max_ratio = vars[2], ignore = vars[3];
    for (int32_t y = 0; y < cnt_y; ++y)
        for (int32_t x = 0; x < cnt_x; ++x) {
            const size_t t = (size_t)y * cnt_x + x, xm = (size_t)x * cnt_y + y;
            if (blockage[xm] >= ignore) actionable[t] = 0;
            if (!actionable[t]) continue;
            const float ratio = orc_rudy_congestion(demand[xm], blockage[xm]);
            if (min_cong > 0.0f && ratio >= min_cong) {
                const float r = powf(ratio / min_cong, coef);
                inflation_ratio[t] = r < max_ratio ? r : max_ratio;
            }
        }
    double total = 0.0;
    int32_t over = 0, m = 0;
    double* cong = (double*)malloc(sizeof(double) * (n ? n : 1));
    for (int32_t y = 0; y < cnt_y; ++y)
        for (int32_t x = 0; x < cnt_x; ++x) {
            const size_t t = (size_t)y * cnt_x + x, xm = (size_t)x * cnt_y + y;
            if (!actionable[t]) continue;
            const float ratio = orc_rudy_congestion(demand[xm], blockage[xm]);
            if (isfinite(ratio) && ratio >= 0.0f) {
                const double r = (double)ratio;
                total += (-1.0 + r) > 0.0 ? (-1.0 + r) : 0.0;
                cong[m++] = r;
                if (ratio > 1.0f) ++over;
            }
        }
    stats[0] = total;
    stats[1] = over;
    stats[2] = m;
    double top[4] = {0.0, 0.0, 0.0, 0.0}, rc = 0.0;
    if (m > 0) {
        qsort(cong, (size_t)m, sizeof(double), desc);
        orc_top_averages(m, cong, top);
        const double k0 = vars[4], k1 = vars[5], k2 = vars[6], k3 = vars[7];
        rc = (double)(float)((k0 * top[0] + k1 * top[1] + k2 * top[2] + k3 * top[3]) / (double)(float)(vars[4] + vars[5] + vars[6] + vars[7]));
    }
    for (int s = 0; s < 4; ++s) stats[3 + s] = top[s];
    stats[7] = rc;
    free(cong);
}

/* set_size (nesterov.rs:3754-3760): the box rebuilt from halves around its centre. */
void orc_set_size(const int32_t* box, int32_t dx, int32_t dy, int32_t* out) {
    const int32_t cx = (box[0] + box[2]) / 2, cy = (box[1] + box[3]) / 2;
    out[0] = cx - dx / 2;
    out[1] = cy - dy / 2;
    out[2] = cx + dx / 2;
    out[3] = cy + dy / 2;
}

/* RouteBase::routability's per-cell loop (nesterov.rs:3945-3960): each standard cell (std[i]) whose
 * density centre (centres, 2 per cell) lies in a tile (y-major, origin (0, 0), tile size `tile`,
 * the last column and row clamping) of inflated ratio r > 1 takes round(d * sqrt r) per side,
 * set_size around its own box's centre (boxes, 4 per cell, updated); returns the area added. */
int64_t orc_route_inflate(int32_t n, int32_t* boxes, const int32_t* centres, const uint8_t* std_cell, int32_t cnt_x, int32_t cnt_y, int32_t tile,
                          const float* ratio) {
    int64_t delta = 0;
    for (int32_t i = 0; i < n; ++i) {
        if (!std_cell[i]) continue;
        int32_t ix = centres[2 * i] / tile, iy = centres[2 * i + 1] / tile;
        if (ix > cnt_x - 1) ix = cnt_x - 1;
        if (iy > cnt_y - 1) iy = cnt_y - 1;
        if (ix < 0 || iy < 0) continue;
        const float r = ratio[(size_t)iy * cnt_x + ix];
        if ((double)r <= 1.0) continue;
        int32_t* b = boxes + 4 * (size_t)i;
        const int32_t dx = b[2] - b[0], dy = b[3] - b[1];
        const int64_t prev = (int64_t)dx * dy;
        const float root = sqrtf(r);
        const int32_t nx = (int32_t)roundf((float)dx * root), ny = (int32_t)roundf((float)dy * root);
        int32_t nb[4];
        orc_set_size(b, nx, ny, nb);
        memcpy(b, nb, sizeof nb);
        delta += (int64_t)(nb[2] - nb[0]) * (nb[3] - nb[1]) - prev;
    }
    return delta;
}

/* updateDensitySize (nesterov.rs:2022-2036) in the reference's extended precision: a side below
 * sqrt(2) bins spreads to sqrt(2) bins (the size rounded from the float bin size) with the scale
 * d / thr; out: 2 per cell the density size, scale per cell. */
void orc_density_size(int32_t n, const int32_t* sizes, double bin_x, double bin_y, int32_t* dsize, float* scale) {
    const long double s2 = 1.414213562373095048801L;
    const long double thr_x = s2 * (long double)bin_x, thr_y = s2 * (long double)bin_y;
    const float size_x = (float)(s2 * (long double)(float)bin_x), size_y = (float)(s2 * (long double)(float)bin_y);
    for (int32_t i = 0; i < n; ++i) {
        const int32_t dx = sizes[2 * i], dy = sizes[2 * i + 1];
        float sx = 1.0f, sy = 1.0f, dsx = (float)dx, dsy = (float)dy;
        if ((long double)dx < thr_x) {
            sx = (float)dx / (float)thr_x;
            dsx = size_x;
        }
        if ((long double)dy < thr_y) {
            sy = (float)dy / (float)thr_y;
            dsy = size_y;
        }
        dsize[2 * i] = (int32_t)dsx;
        dsize[2 * i + 1] = (int32_t)dsy;
        scale[i] = sx * sy;
    }
}
