// Routability's congestion measurement on the device (route_base.rs): RUDY tile demand from every
// net's terminal box (Rudy::calculate / visitNetTiles, 191-253), the tiles' blockage from the
// router's capacity reductions (getResourceReductions, 183-188), the tile grid's actionable tiles
// and inflation ratios (calculateRudyTiles, 443-488) and the top 0.5/1/2/5 % averages with the
// weighted congestion (updateRudyAverage, 493-536).
//
// The reference sums each tile's demand in float in the nets' order.  Here every (tile, net)
// contribution is formed in parallel with the reference's float expressions, the pairs are radix
// sorted by (tile, net in the problem's order), and one lane per tile adds its run in that order:
// the same floats, bit for bit.  The averages take the actionable tiles' congestions sorted in
// descending order and one lane sums each share in that order, FP64 as the reference does; the
// total overflow likewise in the tiles' order.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <climits>
#include <cmath>

#include "kernels.h"

namespace gplk {

namespace {

constexpr int kRB = 256;

// makeGrid's tile (x, y): from the die's lower left; the last column and row stretched by what
// the grid leaves of the die, measured from the origin.
__device__ __forceinline__ int4 rudy_rect(const RudyGrid& g, int x, int y) {
    const int cx = g.lx + x * g.tile, cy = g.ly + y * g.tile;
    return make_int4(cx, cy, cx + g.tile + (x == g.cnt_x - 1 ? g.ux - g.cnt_x * g.tile : 0),
                     cy + g.tile + (y == g.cnt_y - 1 ? g.uy - g.cnt_y * g.tile : 0));
}

// visitNetTiles' range of a box: false for an inverted box or one of no area.
__device__ __forceinline__ bool rudy_range(const RudyGrid& g, int4 b, int& x0, int& x1, int& y0, int& y1, float& cong) {
    if (b.x > b.z || b.y > b.w) return false;
    const long long area = (long long)(b.z - b.x) * (long long)(b.w - b.y);
    if (area == 0) return false;
    const float hpwl = (float)((b.z - b.x) + (b.w - b.y));
    cong = (hpwl * (float)g.wire_width) / (float)area;
    x0 = max(0, (b.x - g.lx) / g.tile);
    x1 = min(g.cnt_x - 1, (b.z - g.lx) / g.tile);
    y0 = max(0, (b.y - g.ly) / g.tile);
    y1 = min(g.cnt_y - 1, (b.w - g.ly) / g.tile);
    return x0 <= x1 && y0 <= y1;
}

__global__ void k_rudy_count(int E, const NetRec* rec, RudyGrid g, unsigned long long* count) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) {
        int x0, x1, y0, y1;
        float cong;
        count[e] = rudy_range(g, rec[e].box, x0, x1, y0, y1, cong) ? (unsigned long long)(x1 - x0 + 1) * (unsigned long long)(y1 - y0 + 1) : 0ull;
    }
}

// Each net's pairs at its offset: key (tile x-major) << 32 | the net's index in the problem's
// order, value its demand in the tile (route_base.rs:240-251); tiles the box does not strictly
// overlap get the key ~0 (sorted last, not summed).  Nets of many tiles take a block each.
__device__ __forceinline__ void rudy_emit_one(const RudyGrid& g, int4 b, int x0, int y0, int ny, float cong, unsigned long long o,
                                              unsigned long long i, unsigned orig, unsigned long long* keys, float* vals) {
    const int x = x0 + (int)(i / (unsigned long long)ny), y = y0 + (int)(i % (unsigned long long)ny);
    const int4 r = rudy_rect(g, x, y);
    unsigned long long key = ~0ull;
    float v = 0.0f;
    if (r.z > b.x && r.x < b.z && r.w > b.y && r.y < b.w) {
        const long long inter = (long long)(min(b.z, r.z) - max(b.x, r.x)) * (long long)(min(b.w, r.w) - max(b.y, r.y));
        const long long tile_area = (long long)(r.z - r.x) * (long long)(r.w - r.y);
        const float ratio = (float)inter / (float)tile_area;
        v = cong * ratio * 100.0f;
        key = ((unsigned long long)((unsigned)x * (unsigned)g.cnt_y + (unsigned)y) << 32) | orig;
    }
    keys[o + i] = key;
    vals[o + i] = v;
}

constexpr unsigned long long kRudyBig = 4096;  // tiles above which a net's pairs take a block

__global__ void k_rudy_emit(int E, const NetRec* rec, RudyGrid g, const unsigned long long* offset, const int32_t* net_orig,
                            unsigned long long* keys, float* vals) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) {
        const int4 b = rec[e].box;
        int x0, x1, y0, y1;
        float cong;
        if (!rudy_range(g, b, x0, x1, y0, y1, cong)) continue;
        const unsigned long long cnt = (unsigned long long)(x1 - x0 + 1) * (unsigned long long)(y1 - y0 + 1);
        if (cnt > kRudyBig) continue;  // k_rudy_emit_big
        for (unsigned long long i = 0; i < cnt; ++i) rudy_emit_one(g, b, x0, y0, y1 - y0 + 1, cong, offset[e], i, (unsigned)net_orig[e], keys, vals);
    }
}

__global__ void k_rudy_emit_big(int E, const NetRec* rec, RudyGrid g, const unsigned long long* offset, const int32_t* net_orig,
                                unsigned long long* keys, float* vals) {
    for (int e = blockIdx.x; e < E; e += gridDim.x) {
        const int4 b = rec[e].box;
        int x0, x1, y0, y1;
        float cong;
        if (!rudy_range(g, b, x0, x1, y0, y1, cong)) continue;
        const unsigned long long cnt = (unsigned long long)(x1 - x0 + 1) * (unsigned long long)(y1 - y0 + 1);
        if (cnt <= kRudyBig) continue;
        for (unsigned long long i = threadIdx.x; i < cnt; i += blockDim.x)
            rudy_emit_one(g, b, x0, y0, y1 - y0 + 1, cong, offset[e], i, (unsigned)net_orig[e], keys, vals);
    }
}

// One lane per run of equal tiles in the sorted pairs adds the run in order (the nets' order).
__global__ void k_rudy_sum(unsigned long long total, const unsigned long long* keys, const float* vals, float* demand) {
    for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < total;
         i += (unsigned long long)gridDim.x * blockDim.x) {
        const unsigned long long k = keys[i];
        if (k == ~0ull) continue;
        const unsigned t = (unsigned)(k >> 32);
        if (i > 0 && (unsigned)(keys[i - 1] >> 32) == t && keys[i - 1] != ~0ull) continue;
        float s = 0.0f;
        for (unsigned long long j = i; j < total && keys[j] != ~0ull && (unsigned)(keys[j] >> 32) == t; ++j) s += vals[j];
        demand[t] = s;
    }
}

__global__ void k_rudy_blockage(unsigned n, const uint8_t* cap_red, float* demand, float* blockage) {
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        demand[i] = 0.0f;
        float b = 0.0f;
        if (cap_red) {
            const uint8_t cap = cap_red[2 * i], red = cap_red[2 * i + 1];
            b = cap == 0 ? 1.0f : fminf((float)red / (float)cap, 1.0f);
        }
        blockage[i] = b;
    }
}

__device__ __forceinline__ float rudy_congestion(float demand, float blockage) {
    const float available = 1.0f - blockage;
    if (available <= 0.0f) return INFINITY;
    return (demand / 100.0f) / available;
}

// markTilesWithMovableArea: the tile (origin 0, 0; the last column and row clamp) of every standard
// instance cell's density centre.
__global__ void k_rudy_mark(int n, CellsDev c, RudyGrid g, uint8_t* act) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint8_t f = c.flags[k];
        if (!(f & 1u) || (f & 2u)) continue;  // GPL_CELL_INSTANCE, not GPL_CELL_MACRO
        const int2 m = c.ctr[k];
        const int ix = min(m.x / g.tile, g.cnt_x - 1), iy = min(m.y / g.tile, g.cnt_y - 1);
        if (ix < 0 || iy < 0) continue;
        act[(size_t)iy * g.cnt_x + ix] = 1;
    }
}

// Per tile (y-major): the blocked ones unmarked, an actionable tile's inflation ratio
// (updateTileInflationRatio) and its congestion for the averages (-1: not counted).
__global__ void k_rudy_ratio(RudyGrid g, RouteVarsDev v, const float* demand, const float* blockage, uint8_t* act, float* ratio,
                             float* cong) {
    const unsigned n = (unsigned)g.cnt_x * (unsigned)g.cnt_y;
    for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const unsigned y = t / (unsigned)g.cnt_x, x = t - y * (unsigned)g.cnt_x;
        const unsigned xm = x * (unsigned)g.cnt_y + y;
        float r = 1.0f, cg = -1.0f;
        bool a = act[t] != 0;
        if (blockage[xm] >= v.ignore_edge_ratio) a = false;
        if (a) {
            const float c = rudy_congestion(demand[xm], blockage[xm]);
            if (v.min_congestion > 0.0f && c >= v.min_congestion) {
                const float q = c / v.min_congestion;
                const float p = v.coef == 2.0f ? q * q : powf(q, v.coef);  // x * x: x^2 correctly rounded
                r = fminf(p, v.max_ratio);
            }
            if (isfinite(c) && c >= 0.0f) cg = c;
        }
        act[t] = a ? 1 : 0;
        ratio[t] = r;
        cong[t] = cg;
    }
}

// updateRudyAverage's sums in the reference's orders, by one block: the counts from the congestions
// sorted in descending order (the uncounted tiles hold -1 and sort last); the total overflow over the
// overflowed tiles in the tiles' order (a term max(0, c - 1) of any other tile is +0, which leaves a
// non-negative double sum as it is); the shares over the sorted congestions.  The block stages 256
// values at a time in LDS, where its first lane adds them in order.
constexpr int kStatBlock = 256;
__device__ unsigned count_desc_above(const float* sorted, unsigned n, float t, bool or_equal) {
    // the first index whose value is not above t (not at or above t with or_equal), by bisection
    unsigned lo = 0, hi = n;
    while (lo < hi) {
        const unsigned mid = (lo + hi) >> 1;
        const float v = sorted[mid];
        if (or_equal ? v >= t : v > t)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

__global__ void __launch_bounds__(kStatBlock) k_rudy_stats(unsigned n, const float* sorted, const float* over_list, const int* nover,
                                                           RouteVarsDev v, double* out) {
    __shared__ float s_v[kStatBlock];
    __shared__ unsigned s_m, s_over;
    if (threadIdx.x == 0) {
        s_m = count_desc_above(sorted, n, 0.0f, true);
        s_over = (unsigned)*nover;
    }
    __syncthreads();
    const unsigned m = s_m, over = s_over;
    // the total overflow over the overflowed tiles, in the tiles' order
    double total = 0.0;
    for (unsigned b0 = 0; b0 < over; b0 += kStatBlock) {
        const unsigned i = b0 + threadIdx.x;
        s_v[threadIdx.x] = i < over ? over_list[i] : 0.0f;
        __syncthreads();
        if (threadIdx.x == 0) {
            const unsigned e = min(over - b0, (unsigned)kStatBlock);
            for (unsigned j = 0; j < e; ++j) {
                const double r = (double)s_v[j];
                total += (-1.0 + r) > 0.0 ? (-1.0 + r) : 0.0;
            }
        }
        __syncthreads();
    }
    double top[4] = {0.0, 0.0, 0.0, 0.0}, rc = 0.0;
    const double shares[4] = {0.005, 0.01, 0.02, 0.05};
    const double nd = (double)m;
    double sums[4] = {0.0, 0.0, 0.0, 0.0};
    // the elements i < 0.05 m of the sorted list (the widest share: ceil(0.05 m) of them), staged
    // and added in order
    const unsigned lim = m > 0 ? min(m, (unsigned)ceil(shares[3] * nd)) : 0u;
    for (unsigned b0 = 0; b0 < lim; b0 += kStatBlock) {
        const unsigned i = b0 + threadIdx.x;
        s_v[threadIdx.x] = i < lim ? sorted[i] : 0.0f;
        __syncthreads();
        if (threadIdx.x == 0) {
            const unsigned e = min(lim - b0, (unsigned)kStatBlock);
            for (unsigned j = 0; j < e; ++j)
                for (int q = 0; q < 4; ++q)
                    if ((double)(b0 + j) < shares[q] * nd) sums[q] += (double)s_v[j];
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    out[0] = total;
    out[1] = (double)over;
    out[2] = (double)m;
    if (m > 0) {
        top[0] = sums[0] / ceil(0.005 * nd);
        top[1] = sums[1] / ceil(0.010 * nd);
        top[2] = sums[2] / ceil(0.020 * nd);
        top[3] = sums[3] / ceil(0.050 * nd);
        const double k0 = v.rc_k[0], k1 = v.rc_k[1], k2 = v.rc_k[2], k3 = v.rc_k[3];
        const float ks = v.rc_k[0] + v.rc_k[1] + v.rc_k[2] + v.rc_k[3];
        rc = (double)(float)((k0 * top[0] + k1 * top[1] + k2 * top[2] + k3 * top[3]) / (double)ks);
    }
    for (int s = 0; s < 4; ++s) out[3 + s] = top[s];
    out[7] = rc;
}

// The per-cell half of a routability pass (nesterov.rs:3945-3960 with set_size, 3754-3760): a
// standard instance cell whose density centre lies in a tile of inflated ratio r > 1 grows to
// round(d sqrt r) per side, rebuilt from halves around its centre.
__global__ void k_route_inflate(int begin, int n, CellsDev c, RudyGrid g, const float* ratio, int2* sizes, unsigned long long* delta) {
    long long add = 0;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int cell = begin + k;
        const uint8_t f = c.flags[cell];
        if (!(f & 1u) || (f & 2u)) continue;  // GPL_CELL_INSTANCE, not GPL_CELL_MACRO
        const int4 d = c.dbox[cell];
        const int cx = (d.x + d.z) / 2, cy = (d.y + d.w) / 2;
        const int ix = min(cx / g.tile, g.cnt_x - 1), iy = min(cy / g.tile, g.cnt_y - 1);
        if (ix < 0 || iy < 0) continue;
        const float r = ratio[(size_t)iy * g.cnt_x + ix];
        if (!((double)r > 1.0)) continue;
        const int2 s = sizes[k];
        const float root = sqrtf(r);
        const int nx = (int)roundf((float)s.x * root), ny = (int)roundf((float)s.y * root);
        const int2 t = make_int2(2 * (nx / 2), 2 * (ny / 2));
        add += (long long)t.x * t.y - (long long)s.x * s.y;
        sizes[k] = t;
    }
    for (int o = 32; o > 0; o >>= 1) add += __shfl_xor(add, o, 64);
    if ((threadIdx.x & 63) == 0 && add) atomicAdd(delta, (unsigned long long)add);
}

// updateDensitySize (nesterov.rs:2022-2036): a side narrower than sqrt(2) bins (max_x: the largest
// integer below sqrt(2) size_x, the reference's extended-precision threshold) spreads to
// sqrt(2) bins (size_x, rounded as the reference rounds it) with the scale d / thr; the density box
// rebuilt from halves around its centre; the preconditioner's area the cell's.
__global__ void k_density_size(int begin, int n, CellsDev c, DensitySizeDev ds, const int2* sizes) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int cell = begin + k;
        const int2 s = sizes[k];
        float sx = 1.0f, sy = 1.0f, dsx = (float)s.x, dsy = (float)s.y;
        if (s.x <= ds.max_x) {
            sx = (float)s.x / ds.thr_x;
            dsx = ds.size_x;
        }
        if (s.y <= ds.max_y) {
            sy = (float)s.y / ds.thr_y;
            dsy = ds.size_y;
        }
        const int ddx = (int)dsx, ddy = (int)dsy;
        const int4 d = c.dbox[cell];
        const int cx = (d.x + d.z) / 2, cy = (d.y + d.w) / 2;
        c.dbox[cell] = make_int4(cx - ddx / 2, cy - ddy / 2, cx + ddx / 2, cy + ddy / 2);
        c.dscale[cell] = sx * sy;
        const float2 p = c.pre[cell];
        c.pre[cell] = make_float2(p.x, (float)s.x * (float)s.y);
    }
}

__global__ void k_park(int k, const int32_t* pos, const int32_t* idx, ParkSlots v, float2* store, int stride, int restore) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < k; i += gridDim.x * blockDim.x) {
        const int p = pos[i];
        const size_t q = (size_t)idx[i];
        for (int s = 0; s < v.count; ++s) {
            if (restore)
                v.v[s][p] = store[(size_t)s * stride + q];
            else
                store[(size_t)s * stride + q] = v.v[s][p];
        }
    }
}

__global__ void k_flag_list(int k, const int32_t* pos, uint8_t* flags, uint8_t set, uint8_t clear) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < k; i += gridDim.x * blockDim.x) {
        const int p = pos[i];
        flags[p] = (uint8_t)((flags[p] & (uint8_t)~clear) | set);
    }
}

__global__ void k_non_place(int m, const long long* raw, const long long* uni, float td, long long* np, float2* npf) {
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < m; b += gridDim.x * blockDim.x) {
        const long long r = raw[b], u = uni[b];
        long long v = r != 0 ? (long long)((float)r * td) : 0;
        long long vu = 0;
        if (u >= 0) {
            vu = (long long)((float)u * td);
            const long long cap = (long long)((float)u * td * 1.10f);
            if (v > cap) v = cap;
        }
        np[b] = v;
        npf[b] = make_float2((float)v, (float)vu);
    }
}

// The congested tiles (inflation ratio above 1) listed per column, rows ascending: the counts,
// then (after a scan into col_start) the rows.
__global__ void k_col_count(RudyGrid g, const float* ratio, int* cnt) {
    for (int x = blockIdx.x * blockDim.x + threadIdx.x; x < g.cnt_x; x += gridDim.x * blockDim.x) {
        int k = 0;
        for (int y = 0; y < g.cnt_y; ++y) k += ratio[(size_t)y * g.cnt_x + x] > 1.0f ? 1 : 0;
        cnt[x] = k;
    }
}
__global__ void k_col_fill(RudyGrid g, const float* ratio, const int* start, int* col_y) {
    for (int x = blockIdx.x * blockDim.x + threadIdx.x; x < g.cnt_x; x += gridDim.x * blockDim.x) {
        int at = start[x];
        for (int y = 0; y < g.cnt_y; ++y)
            if (ratio[(size_t)y * g.cnt_x + x] > 1.0f) col_y[at++] = y;
    }
}

// Each net's demand over the congested tiles of its box, in visitNetTiles' order (x outer, y inner):
// per column the first listed row at or above the box's by bisection, then the rows up to its top.
__global__ void k_net_congested_demand(int E, const NetRec* rec, const int32_t* net_orig, RudyGrid g, const int* col_start, const int* col_y,
                                       float* out) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) {
        const int4 b = rec[e].box;
        int x0, x1, y0, y1;
        float cong;
        float d = 0.0f;
        if (rudy_range(g, b, x0, x1, y0, y1, cong)) {
            for (int x = x0; x <= x1; ++x) {
                int lo = col_start[x], hi = col_start[x + 1];
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (col_y[mid] < y0)
                        lo = mid + 1;
                    else
                        hi = mid;
                }
                for (int q = lo; q < col_start[x + 1]; ++q) {
                    const int y = col_y[q];
                    if (y > y1) break;
                    const int4 r = rudy_rect(g, x, y);
                    if (!(r.z > b.x && r.x < b.z && r.w > b.y && r.y < b.w)) continue;
                    const long long inter = (long long)(min(b.z, r.z) - max(b.x, r.x)) * (long long)(min(b.w, r.w) - max(b.y, r.y));
                    const long long tile_area = (long long)(r.z - r.x) * (long long)(r.w - r.y);
                    const float rt = (float)inter / (float)tile_area;
                    d += cong * rt * 100.0f;
                }
            }
        }
        out[net_orig[e]] = d;
    }
}

int grid_of(unsigned long long n) { return (int)std::max<unsigned long long>(1, std::min<unsigned long long>((n + kRB - 1) / kRB, 8192)); }

template <typename T>
void grow(T*& p, size_t& cap, size_t n) {
    if (n <= cap && p) return;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = std::max<size_t>(n, 1);
    if (hipMalloc(&p, sizeof(T) * cap) != hipSuccess) {
        p = nullptr;
        cap = 0;
    }
}

}  // namespace

void RudyBufs::release() {
    if (count) (void)hipFree(count);
    if (offset) (void)hipFree(offset);
    if (keys) (void)hipFree(keys);
    if (keys2) (void)hipFree(keys2);
    if (vals) (void)hipFree(vals);
    if (vals2) (void)hipFree(vals2);
    if (temp) (void)hipFree(temp);
    if (demand) (void)hipFree(demand);
    if (blockage) (void)hipFree(blockage);
    if (act) (void)hipFree(act);
    if (ratio) (void)hipFree(ratio);
    if (cong) (void)hipFree(cong);
    if (sorted) (void)hipFree(sorted);
    if (stats) (void)hipFree(stats);
    if (over) (void)hipFree(over);
    if (nover) (void)hipFree(nover);
    if (col_start) (void)hipFree(col_start);
    if (col_y) (void)hipFree(col_y);
    *this = RudyBufs{};
}

hipError_t launch_rudy(hipStream_t s, const NetsDev& n, const int32_t* net_orig, const RudyGrid& g, const uint8_t* cap_red_dev, RudyBufs& b) {
    const int E = n.num_nets;
    const size_t tiles = (size_t)g.cnt_x * (size_t)g.cnt_y;
    grow(b.count, b.cap_nets, (size_t)E + 1);
    grow(b.offset, b.cap_off, (size_t)E + 1);
    grow(b.demand, b.cap_tiles, tiles);
    grow(b.blockage, b.cap_tiles2, tiles);
    if (!b.count || !b.offset || !b.demand || !b.blockage) return hipErrorOutOfMemory;
    hipLaunchKernelGGL(k_rudy_blockage, dim3(grid_of(tiles)), dim3(kRB), 0, s, (unsigned)tiles, cap_red_dev, b.demand, b.blockage);
    if (E == 0) return hipGetLastError();
    hipLaunchKernelGGL(k_rudy_count, dim3(grid_of(E)), dim3(kRB), 0, s, E, n.rec, g, b.count);
    size_t tb = 0;
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, tb, b.count, b.offset, E + 1, s);
    if (e != hipSuccess) return e;
    grow(b.temp, b.cap_temp, tb);
    if (!b.temp) return hipErrorOutOfMemory;
    // count[E] is the scan's extra element (the total lands in offset[E])
    e = hipMemsetAsync(b.count + E, 0, sizeof(unsigned long long), s);
    if (e != hipSuccess) return e;
    e = hipcub::DeviceScan::ExclusiveSum(b.temp, tb, b.count, b.offset, E + 1, s);
    if (e != hipSuccess) return e;
    unsigned long long total = 0;
    e = hipMemcpyAsync(&total, b.offset + E, sizeof total, hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    if (total == 0) return hipSuccess;
    if (total > (unsigned long long)INT_MAX) return hipErrorInvalidValue;  // the sort counts in int
    grow(b.keys, b.cap_pairs, total);
    grow(b.keys2, b.cap_pairs2, total);
    grow(b.vals, b.cap_pairs3, total);
    grow(b.vals2, b.cap_pairs4, total);
    if (!b.keys || !b.keys2 || !b.vals || !b.vals2) return hipErrorOutOfMemory;
    hipLaunchKernelGGL(k_rudy_emit, dim3(grid_of(E)), dim3(kRB), 0, s, E, n.rec, g, b.offset, net_orig, b.keys, b.vals);
    hipLaunchKernelGGL(k_rudy_emit_big, dim3(std::min(E, 1024)), dim3(kRB), 0, s, E, n.rec, g, b.offset, net_orig, b.keys, b.vals);
    tb = 0;
    e = hipcub::DeviceRadixSort::SortPairs(nullptr, tb, b.keys, b.keys2, b.vals, b.vals2, (int)total, 0, 64, s);
    if (e != hipSuccess) return e;
    grow(b.temp, b.cap_temp, tb);
    if (!b.temp) return hipErrorOutOfMemory;
    e = hipcub::DeviceRadixSort::SortPairs(b.temp, tb, b.keys, b.keys2, b.vals, b.vals2, (int)total, 0, 64, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_rudy_sum, dim3(grid_of(total)), dim3(kRB), 0, s, total, b.keys2, b.vals2, b.demand);
    return hipGetLastError();
}

hipError_t launch_rudy_tiles(hipStream_t s, const CellsDev& c, int ncells, const RudyGrid& g, const RouteVarsDev& v, RudyBufs& b) {
    const size_t tiles = (size_t)g.cnt_x * (size_t)g.cnt_y;
    grow(b.act, b.cap_act, tiles);
    grow(b.ratio, b.cap_ratio, tiles);
    grow(b.cong, b.cap_cong, tiles);
    grow(b.sorted, b.cap_sorted, tiles);
    grow(b.stats, b.cap_stats, 8);
    if (!b.act || !b.ratio || !b.cong || !b.sorted || !b.stats || !b.demand) return hipErrorOutOfMemory;
    hipError_t e = hipMemsetAsync(b.act, 0, tiles, s);
    if (e != hipSuccess) return e;
    if (ncells > 0) hipLaunchKernelGGL(k_rudy_mark, dim3(grid_of(ncells)), dim3(kRB), 0, s, ncells, c, g, b.act);
    hipLaunchKernelGGL(k_rudy_ratio, dim3(grid_of(tiles)), dim3(kRB), 0, s, g, v, b.demand, b.blockage, b.act, b.ratio, b.cong);
    size_t tb = 0;
    e = hipcub::DeviceRadixSort::SortKeysDescending(nullptr, tb, b.cong, b.sorted, (int)tiles, 0, 32, s);
    if (e != hipSuccess) return e;
    grow(b.temp, b.cap_temp, tb);
    if (!b.temp) return hipErrorOutOfMemory;
    e = hipcub::DeviceRadixSort::SortKeysDescending(b.temp, tb, b.cong, b.sorted, (int)tiles, 0, 32, s);
    if (e != hipSuccess) return e;
    // the overflowed tiles' congestions in the tiles' order (the total overflow's terms)
    grow(b.over, b.cap_over, tiles);
    grow(b.nover, b.cap_nover, 1);
    if (!b.over || !b.nover) return hipErrorOutOfMemory;
    auto overflowed = [] __device__(float c) { return c > 1.0f; };
    tb = 0;
    e = hipcub::DeviceSelect::If(nullptr, tb, b.cong, b.over, b.nover, (int)tiles, overflowed, s);
    if (e != hipSuccess) return e;
    grow(b.temp, b.cap_temp, tb);
    if (!b.temp) return hipErrorOutOfMemory;
    e = hipcub::DeviceSelect::If(b.temp, tb, b.cong, b.over, b.nover, (int)tiles, overflowed, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_rudy_stats, dim3(1), dim3(kStatBlock), 0, s, (unsigned)tiles, b.sorted, b.over, b.nover, v, b.stats);
    return hipGetLastError();
}

hipError_t launch_route_inflate(hipStream_t s, int begin, int n, const CellsDev& c, const RudyGrid& g, const float* ratio, int2* sizes,
                                unsigned long long* delta) {
    if (n > 0) hipLaunchKernelGGL(k_route_inflate, dim3(grid_of(n)), dim3(kRB), 0, s, begin, n, c, g, ratio, sizes, delta);
    return hipGetLastError();
}

hipError_t launch_density_size(hipStream_t s, int begin, int n, const CellsDev& c, const DensitySizeDev& d, const int2* sizes) {
    if (n > 0) hipLaunchKernelGGL(k_density_size, dim3(grid_of(n)), dim3(kRB), 0, s, begin, n, c, d, sizes);
    return hipGetLastError();
}

hipError_t launch_park_save(hipStream_t s, int k, const int32_t* pos, const int32_t* idx, const ParkSlots& v, float2* store, int stride) {
    if (k > 0) hipLaunchKernelGGL(k_park, dim3(grid_of(k)), dim3(kRB), 0, s, k, pos, idx, v, store, stride, 0);
    return hipGetLastError();
}

hipError_t launch_park_restore(hipStream_t s, int k, const int32_t* pos, const int32_t* idx, const ParkSlots& v, const float2* store,
                               int stride) {
    if (k > 0) hipLaunchKernelGGL(k_park, dim3(grid_of(k)), dim3(kRB), 0, s, k, pos, idx, v, const_cast<float2*>(store), stride, 1);
    return hipGetLastError();
}

hipError_t launch_flag_list(hipStream_t s, int k, const int32_t* pos, uint8_t* flags, uint8_t set, uint8_t clear) {
    if (k > 0) hipLaunchKernelGGL(k_flag_list, dim3(grid_of(k)), dim3(kRB), 0, s, k, pos, flags, set, clear);
    return hipGetLastError();
}

hipError_t launch_non_place(hipStream_t s, int m, const long long* raw, const long long* uni, float td, long long* non_place,
                            float2* non_place_f) {
    if (m > 0) hipLaunchKernelGGL(k_non_place, dim3(grid_of(m)), dim3(kRB), 0, s, m, raw, uni, td, non_place, non_place_f);
    return hipGetLastError();
}

hipError_t launch_net_congested_demand(hipStream_t s, const NetsDev& n, const int32_t* net_orig, const RudyGrid& g, const float* ratio,
                                       float* out, RudyBufs& b) {
    if (n.num_nets <= 0) return hipSuccess;
    const size_t tiles = (size_t)g.cnt_x * (size_t)g.cnt_y;
    grow(b.col_start, b.cap_col, (size_t)g.cnt_x + 1);
    grow(b.col_y, b.cap_coly, tiles);
    if (!b.col_start || !b.col_y) return hipErrorOutOfMemory;
    hipLaunchKernelGGL(k_col_count, dim3(grid_of(g.cnt_x)), dim3(kRB), 0, s, g, ratio, b.col_start);
    // the count at cnt_x is the scan's extra element (the total lands there)
    hipError_t e = hipMemsetAsync(b.col_start + g.cnt_x, 0, sizeof(int), s);
    if (e != hipSuccess) return e;
    size_t tb = 0;
    e = hipcub::DeviceScan::ExclusiveSum(nullptr, tb, b.col_start, b.col_start, g.cnt_x + 1, s);
    if (e != hipSuccess) return e;
    grow(b.temp, b.cap_temp, tb);
    if (!b.temp) return hipErrorOutOfMemory;
    e = hipcub::DeviceScan::ExclusiveSum(b.temp, tb, b.col_start, b.col_start, g.cnt_x + 1, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_col_fill, dim3(grid_of(g.cnt_x)), dim3(kRB), 0, s, g, ratio, b.col_start, b.col_y);
    hipLaunchKernelGGL(k_net_congested_demand, dim3(grid_of(n.num_nets)), dim3(kRB), 0, s, n.num_nets, n.rec, net_orig, g, b.col_start, b.col_y,
                       out);
    return hipGetLastError();
}

}  // namespace gplk
