// HIP kernels of the ePlace Nesterov loop for CDNA4 (gfx950), 64-wide wavefronts.
//
// Arithmetic follows the reference function it replaces (cited per kernel) with the same float
// types and evaluation order; the build passes -ffp-contract=off so no multiply-add is fused,
// as the reference's is not.  Global sums (overflow, |gradient| sums, the step-length norms)
// reduce in FP64 in a fixed tree (per-block partials, then one block in index order), so every
// run gives the same bits.  Integer work (bin areas, HPWL) is exact: int64 atomics are
// associative.
#include "kernels.h"

#include <cfloat>
#include <climits>
#include <cstring>
#include <type_traits>
#include <vector>

#include <hipcub/hipcub.hpp>

namespace gplk {

constexpr int kBlock = 256;
constexpr int kMaxBlocks = kPartials;

// The dispatcher hands workgroups to the 8 XCDs in turn (workgroup b to XCD b % 8), and each XCD
// has its own L2.  The scatter takes its chunks of Morton-ordered cells by the logical index below,
// which gives XCD x the contiguous range [x G/8, (x+1) G/8) of a G-workgroup grid (G a multiple of
// 8), so that the bins of neighbouring chunks meet in one L2.  The WA kernel gives XCD x the x-th
// of 8 contiguous slices of the nets (the host orders the nets by their first cell in windows of
// 4096 cells, and cuts the slices at equal pin counts), so that the cell centres it gathers and
// the pin gradients it scatters stay within one L2 and the XCDs get equal work.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block(int b, int G) { return (b % kXcds) * (G / kXcds) + b / kXcds; }
__host__ __device__ constexpr int round_xcds(int g) { return (g + kXcds - 1) / kXcds * kXcds; }

// ---------------------------------------------------------------------------------------------
// Shared device helpers
// ---------------------------------------------------------------------------------------------

// nesterov.rs:308 fastExp: (1 + x/1024)^1024, ten squarings.
__device__ __forceinline__ float fast_exp(float x) {
    float e = 1.0f + x / 1024.0f;
#pragma unroll
    for (int i = 0; i < 10; ++i) e *= e;
    return e;
}

// nesterov.rs:275 density_idx_x / 280 density_idx_y: f64 quotient, truncated / ceiled, clamped.
__device__ __forceinline__ void bin_span(int lo, int hi, int org, double size, int cnt, int& a, int& b) {
    int l = (int)((double)(lo - org) / size);
    int u = (int)ceil((double)(hi - org) / size);
    a = l > 0 ? l : 0;
    b = u < cnt ? u : cnt;
}

// Bin edge i along one axis: the grid's origin plus the rounded multiple of the bin size
// (init_bins, nesterov.rs:1879-1881), when the region's tables hold exactly these (edges_rounded);
// else the table's entry.
__device__ __forceinline__ int bin_edge(const int32_t* table, int rounded, int org, double size, int cnt, int i) {
    return rounded ? org + (int)round((double)i * size) : table[min(i, cnt)];
}

// nesterov.rs:298 getOverlapDensityArea.
__device__ __forceinline__ float overlap_area(int blx, int bly, int bux, int buy, int4 d) {
    int lx = max(blx, d.x), ly = max(bly, d.y), ux = min(bux, d.z), uy = min(buy, d.w);
    if (lx >= ux || ly >= uy) return 0.0f;
    return (float)(ux - lx) * (float)(uy - ly);
}

// Rust's f32::max(a, b) for b not NaN: NaN a gives b.
__device__ __forceinline__ float rmax(float a, float b) { return a >= b ? a : b; }

// nesterov.rs:2188 getDensityCoordiLayoutInsideX for a box of half width `half`.
__device__ __forceinline__ float inside(float c, float half, float lo, float hi) {
    float adj = c;
    if (c - half < lo) adj = lo + half;
    if (c + half > hi) adj = hi - half;
    return adj;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Block sum of a double (blockDim.x == kBlock), in a fixed tree; thread 0 gets the result.
__device__ __forceinline__ double block_sum(double v, double* sh) {
    v = wave_sum(v);
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kBlock / 64; ++i) r += sh[i];
    return r;
}

// A cell's density box centred at the truncated point (nesterov.rs:2139 setDensityCenterLocation):
// the halves of the current size, so an odd size loses one.
__device__ __forceinline__ int4 centre_box(int4 d, float x, float y) {
    int cx = (int)x, cy = (int)y;
    int hx = (d.z - d.x) / 2, hy = (d.w - d.y) / 2;
    return make_int4(cx - hx, cy - hy, cx + hx, cy + hy);
}

// ---------------------------------------------------------------------------------------------
// Scatter family
// ---------------------------------------------------------------------------------------------

__global__ void k_zero_bins(unsigned long long* a, unsigned long long* b, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (size_t)gridDim.x * blockDim.x) {
        a[i] = 0ull;
        b[i] = 0ull;
    }
}

// The scatter family: a cell's density box by one of three rules, then its overlaps with the bins
// added into the int64 accumulators (nesterov.rs:2225-2250).
//   PLACE_INIT    initDensity1 (nesterov.rs:2546-2570): the box moved inside the grid
//                 (updateDensityCoordiLayoutInside, 2173), its centre as cur_slp / prev_slp / cur,
//                 the box centred there
//   PLACE_CENTER  updateGCellDensityCenterLocation (nesterov.rs:2268) at coordinates xy
//   PLACE_STORED  the box as k_place_update left it (the trial's Nesterov update runs as its own
//                 kernel so that the WA kernels, which need only the cells' centres, can start
//                 beside this scatter)
// A block takes kChunk consecutive cells (the device keeps each region's cells in Morton order of
// their bins, re-sorted when that order has gone stale, so a chunk touches few distinct bins).  Its
// overlaps are summed in an LDS hash table keyed by bin (kSlots slots, linear probing, keys set
// by compare-and-swap), then each occupied slot is added into the bins with one global atomic;
// an overlap that finds no slot within kProbe probes goes straight to global memory.  The sums
// are of integers: the result does not depend on the path.
enum PlaceMode { PLACE_INIT = 0, PLACE_CENTER = 1, PLACE_STORED = 2 };
constexpr int kSlots = 2048;
constexpr int kProbe = 16;
constexpr unsigned kEmpty = 0xffffffffu;

struct PlaceArgs {
    const float2* cur_slp;  // UPDATE: read; INIT: written
    const float2* cur_sum;
    const float2* cur;      // UPDATE: read; INIT: written (as cur)
    float2* next;           // UPDATE: written; INIT: prev_slp written
    float2* next_slp;       // UPDATE: written; INIT: cur_slp written
    float2* cur_out;        // INIT: cur written
    const float2* xy;       // CENTER
    float s, coeff;
    unsigned long long* global_adds;  // count of the atomic adds sent to HBM
    unsigned long long* zero;         // cleared by the first thread (the trial's HPWL counter)
};

// Adds v into the LDS table's slot for bin k; false when no slot was found.
template <typename V>
__device__ __forceinline__ bool table_add(unsigned* keys, V* vals, unsigned k, V v) {
    unsigned slot = (k * 2654435761u) >> 21;  // 11 bits: kSlots = 2048
#pragma unroll 1
    for (int probe = 0; probe < kProbe; ++probe) {
        unsigned cur = keys[slot];
        if (cur == kEmpty) {
            const unsigned old = atomicCAS(&keys[slot], kEmpty, k);
            cur = old == kEmpty ? k : old;
        }
        if (cur == k) {
            atomicAdd(&vals[slot], v);
            return true;
        }
        slot = (slot + 1) & (kSlots - 1);
    }
    return false;
}

// nesterovUpdateCoordinates (nesterov.rs:2718-2745) for one unlocked cell: next and next_slp from
// the step and the momentum, each kept inside the region (`inside`, 2188-2213) by the box's half
// size.
__device__ __forceinline__ void nesterov_point(float2 a, float2 sm, float2 cu, float s, float coeff, int4 d, float lx, float ly,
                                               float ux, float uy, float2& nxt, float2& nsl) {
    const float nx = a.x + s * sm.x;
    const float ny = a.y + s * sm.y;
    const float sx = nx + coeff * (nx - cu.x);
    const float sy = ny + coeff * (ny - cu.y);
    const float hx = (float)(d.z - d.x) / 2.0f, hy = (float)(d.w - d.y) / 2.0f;
    nxt = make_float2(inside(nx, hx, lx, ux), inside(ny, hy, ly, uy));
    nsl = make_float2(inside(sx, hx, lx, ux), inside(sy, hy, ly, uy));
}

// The trial's placement step for every cell of a region: the Nesterov update (a locked cell,
// flag 4, holds: next = cur, next_slp = cur_slp), the density box centred at next_slp
// (updateGCellDensityCenterLocation, nesterov.rs:2139-2153) and the cell's centre, where its pins
// sit at centre + offset (updatePinLocation, 2268-2276) for the WA kernels.  The first thread
// also clears the trial's HPWL counter.
__device__ __forceinline__ void place_update_body(RegionDev r, CellsDev c, PlaceArgs pa, const unsigned bid_, const unsigned nblk_) {
    int n = r.cell_end - r.cell_begin;
    if (r.ctl) {
        // a device-driven trial: the step, the momentum and the rotated slots from the control
        // record's line (read at once; a skipped region has no cells to move, so the loads are not
        // held back behind a branch)
        const CtlHot h = r.ctl->hot;
        pa.cur_slp = h.cur_slp;
        pa.cur_sum = h.cur_sum;
        pa.cur = h.cur;
        pa.next = h.next;
        pa.next_slp = h.next_slp;
        pa.s = h.step_length;
        pa.coeff = h.coeff;
        pa.zero = nullptr;
        if (h.flags & 1) n = 0;
    }
    const float lx = (float)r.lx, ly = (float)r.ly, ux = (float)r.ux, uy = (float)r.uy;
    if (pa.zero && bid_ == 0 && threadIdx.x == 0) *pa.zero = 0ull;
    for (int k = bid_ * blockDim.x + threadIdx.x; k < n; k += nblk_ * blockDim.x) {
        const int g = r.cell_begin + k;
        int4 d = c.dbox[g];
        const uint8_t f = c.flags[g];
        const float2 a = pa.cur_slp[k], sm = pa.cur_sum[k], cu = pa.cur[k];
        float2 nxt, nsl;
        nesterov_point(a, sm, cu, pa.s, pa.coeff, d, lx, ly, ux, uy, nxt, nsl);
        if (f & 4u) {
            nxt = cu;
            nsl = a;
        }
        pa.next[k] = nxt;
        pa.next_slp[k] = nsl;
        d = centre_box(d, nsl.x, nsl.y);
        c.dbox[g] = d;
        c.ctr[g] = make_int2((d.x + d.z) / 2, (d.y + d.w) / 2);
    }
}

__global__ void __launch_bounds__(kBlock) k_place_update(RegionDev r, CellsDev c, PlaceArgs pa) {
    place_update_body(r, c, pa, blockIdx.x, gridDim.x);
}

// V: the LDS tables' sum type.  32 bits when a chunk's sum in one bin cannot reach 2^32 (every
// overlap, times its scale and target, is at most the bin's area, so kChunk * the largest bin area
// bounds it; the host checks that), which halves the tables and doubles the blocks per CU; else 64.
//
// A chunk first finds the box of bins its cells cover.  When that box holds at most kDense bins
// (cells in Morton order: most chunks), the overlaps are added into a dense LDS tile indexed by
// the bin's place in the box, one LDS atomic per overlap.  When it does not, but the box of the
// chunk's middle half (its cells 128-383 in the storage order) does, the tile covers that box,
// widened evenly while it fits, and the overlaps outside it go straight to the bins in global
// memory: between two re-sorts the storage order decays as cells move, and a few far cells
// would otherwise send the whole chunk to the hash table.  Otherwise the hash table.
// CH: the cells a block takes at a time (a tile shape the autotuner picks per region: 256 or 512).
template <int MODE, typename V, int CH>
__device__ __forceinline__ void place_scatter_body(RegionDev r, CellsDev c, PlaceArgs pa, const unsigned bid_, const unsigned nblk_) {
    if (r.ctl && (r.ctl->hot.flags & 1)) return;
    constexpr int kChunk = CH;
    constexpr int kCellsPerThread = CH / kBlock;
    // the trial's critical chain (scatter -> bins -> solve) shares the CUs with the WA kernels on
    // the side streams: its waves take the SIMDs' issue slots first
    __builtin_amdgcn_s_setprio(2);
    constexpr int kDense = sizeof(V) == 4 ? 3072 : 2048;
    constexpr int kHashBytes = kSlots * (int)(sizeof(unsigned) + 2 * sizeof(V));
    constexpr int kDenseBytes = 2 * kDense * (int)sizeof(V);
    __shared__ __align__(16) unsigned char lds[kHashBytes > kDenseBytes ? kHashBytes : kDenseBytes];
    __shared__ int s_box[8][kBlock / 64];
    __shared__ unsigned s_adds;
    unsigned* t_key = reinterpret_cast<unsigned*>(lds);
    V* t_inst = reinterpret_cast<V*>(lds + kSlots * sizeof(unsigned));
    V* t_fill = t_inst + kSlots;
    V* d_inst = reinterpret_cast<V*>(lds);
    V* d_fill = d_inst + kDense;
    const int n = r.cell_end - r.cell_begin;
    const float lx = (float)r.lx, ly = (float)r.ly, ux = (float)r.ux, uy = (float)r.uy;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (threadIdx.x == 0) s_adds = 0;
    if (pa.zero && bid_ == 0 && threadIdx.x == 0) *pa.zero = 0ull;
    const int chunks = (n + kChunk - 1) / kChunk;
    const int per = (chunks + nblk_ - 1) / nblk_;
    const int L = xcd_block(bid_, nblk_);
    const int ch_end = min((L + 1) * per, chunks);
    for (int ch = L * per; ch < ch_end; ++ch) {
        const int base = ch * kChunk;
        int4 dd[kCellsPerThread];
        int x0[kCellsPerThread], x1[kCellsPerThread], y0[kCellsPerThread], y1[kCellsPerThread];
        float scv[kCellsPerThread];
        uint8_t fl[kCellsPerThread];
        int bx0 = INT_MAX, by0 = INT_MAX, bx1 = INT_MIN, by1 = INT_MIN;
        int cx0 = INT_MAX, cy0 = INT_MAX, cx1 = INT_MIN, cy1 = INT_MIN;  // the chunk's middle half
#pragma unroll
        for (int j = 0; j < kCellsPerThread; ++j) {
            const int k = base + j * kBlock + threadIdx.x;
            x0[j] = x1[j] = y0[j] = y1[j] = 0;
            fl[j] = 0;
            scv[j] = 0.0f;
            dd[j] = make_int4(0, 0, 0, 0);
            if (k >= n) continue;
            const int g = r.cell_begin + k;
            int4 d = c.dbox[g];
            fl[j] = c.flags[g];
            scv[j] = c.dscale[g];
            float2 ns = make_float2(0.0f, 0.0f);
            if (MODE == PLACE_INIT) {
                const int ddx = d.z - d.x, ddy = d.w - d.y;
                float tlx = rmax((float)d.x, lx), tly = rmax((float)d.y, ly);
                if (tlx + (float)ddx > ux) tlx = (float)(r.ux - ddx);
                if (tly + (float)ddy > uy) tly = (float)(r.uy - ddy);
                const int ilx = (int)tlx, ily = (int)tly;
                d = make_int4(ilx, ily, ilx + ddx, ily + ddy);
                ns = make_float2((float)((d.z + d.x) / 2), (float)((d.w + d.y) / 2));
                pa.next_slp[k] = ns;  // cur_slp
                pa.next[k] = ns;      // prev_slp
                pa.cur_out[k] = ns;   // cur
            } else if (MODE == PLACE_CENTER) {
                ns = pa.xy[k];
            }
            if (MODE != PLACE_STORED) {
                d = centre_box(d, ns.x, ns.y);
                c.dbox[g] = d;
                // the cell's centre; its pins sit at centre + offset (updatePinLocation,
                // nesterov.rs:2268-2276), which the WA kernels read from here
                c.ctr[g] = make_int2((d.x + d.z) / 2, (d.y + d.w) / 2);
            }
            dd[j] = d;
            bin_span(d.x, d.z, r.lx, r.size_x, r.cnt_x, x0[j], x1[j]);
            bin_span(d.y, d.w, r.ly, r.size_y, r.cnt_y, y0[j], y1[j]);
            if (x1[j] > x0[j] && y1[j] > y0[j]) {
                bx0 = min(bx0, x0[j]);
                by0 = min(by0, y0[j]);
                bx1 = max(bx1, x1[j]);
                by1 = max(by1, y1[j]);
                const int q = j * kBlock + (int)threadIdx.x;
                if (q >= kChunk / 4 && q < 3 * kChunk / 4) {
                    cx0 = min(cx0, x0[j]);
                    cy0 = min(cy0, y0[j]);
                    cx1 = max(cx1, x1[j]);
                    cy1 = max(cy1, y1[j]);
                }
            }
        }
        // the chunk's box of bins
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            bx0 = min(bx0, __shfl_xor(bx0, o, 64));
            by0 = min(by0, __shfl_xor(by0, o, 64));
            bx1 = max(bx1, __shfl_xor(bx1, o, 64));
            by1 = max(by1, __shfl_xor(by1, o, 64));
            cx0 = min(cx0, __shfl_xor(cx0, o, 64));
            cy0 = min(cy0, __shfl_xor(cy0, o, 64));
            cx1 = max(cx1, __shfl_xor(cx1, o, 64));
            cy1 = max(cy1, __shfl_xor(cy1, o, 64));
        }
        if (lane == 0) {
            s_box[0][w] = bx0;
            s_box[1][w] = by0;
            s_box[2][w] = bx1;
            s_box[3][w] = by1;
            s_box[4][w] = cx0;
            s_box[5][w] = cy0;
            s_box[6][w] = cx1;
            s_box[7][w] = cy1;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kBlock / 64; ++i) {
            bx0 = min(bx0, s_box[0][i]);
            by0 = min(by0, s_box[1][i]);
            bx1 = max(bx1, s_box[2][i]);
            by1 = max(by1, s_box[3][i]);
            cx0 = min(cx0, s_box[4][i]);
            cy0 = min(cy0, s_box[5][i]);
            cx1 = max(cx1, s_box[6][i]);
            cy1 = max(cy1, s_box[7][i]);
        }
        int BW = bx1 - bx0, BH = by1 - by0;
        bool dense = BW > 0 && BH > 0 && BW * BH <= kDense;
        if (!dense && cx1 > cx0 && cy1 > cy0 && (cx1 - cx0) * (cy1 - cy0) <= kDense) {
            // the tile over the middle half's box, widened evenly within the chunk's box while it
            // fits; overlaps outside it go straight to the bins in global memory
            int m = 0;
            while (m < 64 && (cx1 - cx0 + 2 * (m + 1)) * (cy1 - cy0 + 2 * (m + 1)) <= kDense) ++m;
            bx0 = max(bx0, cx0 - m);
            by0 = max(by0, cy0 - m);
            bx1 = min(bx1, cx1 + m);
            by1 = min(by1, cy1 + m);
            BW = bx1 - bx0;
            BH = by1 - by0;
            dense = true;
        }
        const int nd = dense ? BW * BH : 0;
        if (dense) {
            for (int i = threadIdx.x; i < nd; i += kBlock) {
                d_inst[i] = 0;
                d_fill[i] = 0;
            }
        } else {
            for (int i = threadIdx.x; i < kSlots; i += kBlock) {
                t_key[i] = kEmpty;
                t_inst[i] = 0;
                t_fill[i] = 0;
            }
        }
        __syncthreads();
        unsigned adds = 0;
#pragma unroll
        for (int j = 0; j < kCellsPerThread; ++j) {
            if (x1[j] <= x0[j] || y1[j] <= y0[j]) continue;
            const int4 d = dd[j];
            const bool inst = fl[j] & 1u, macro = inst && (fl[j] & 2u);
            const float sc = scv[j];
            V* vals = dense ? (inst ? d_inst : d_fill) : (inst ? t_inst : t_fill);
            unsigned long long* glob = inst ? r.inst : r.filler;
            auto add = [&](int x, int y, int ex0, int ey0, int ex1, int ey1) {
                float a = overlap_area(ex0, ey0, ex1, ey1, d) * sc;
                if (macro) a *= r.target_density;
                // the reference's `as i64` (truncation); with 32-bit tables every overlap is below
                // 2^32 (at most its bin's area), so the one-instruction unsigned conversion gives
                // the same integer
                const long long v = sizeof(V) == 4 ? (long long)(unsigned)a : (long long)a;
                if (v == 0) return;
                if (dense) {
                    const unsigned ux = (unsigned)(x - bx0), uy = (unsigned)(y - by0);
                    if (ux < (unsigned)BW && uy < (unsigned)BH) {
                        atomicAdd(&vals[uy * (unsigned)BW + ux], (V)v);
                    } else {
                        atomicAdd(glob + ((unsigned)y * (unsigned)r.cnt_x + (unsigned)x), (unsigned long long)v);
                        ++adds;
                    }
                    return;
                }
                const unsigned kb = (unsigned)y * (unsigned)r.cnt_x + (unsigned)x;
                if (!table_add<V>(t_key, vals, kb, (V)v)) {
                    atomicAdd(glob + kb, (unsigned long long)v);
                    ++adds;
                }
            };
            if (x1[j] - x0[j] <= 3 && y1[j] - y0[j] <= 3) {
                // the edges of a 3 x 3 bin window read at once (indices clamped, no guards, so
                // that the compiler waits once): the span of any cell up to twice a bin's size
                int ex[4], ey[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    ex[i] = r.bin_x[min(x0[j] + i, r.cnt_x)];
                    ey[i] = r.bin_y[min(y0[j] + i, r.cnt_y)];
                }
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    if (q >= y1[j] - y0[j]) break;
#pragma unroll
                    for (int i = 0; i < 3; ++i) {
                        if (i >= x1[j] - x0[j]) break;
                        add(x0[j] + i, y0[j] + q, ex[i], ey[q], ex[i + 1], ey[q + 1]);
                    }
                }
            } else {
                for (int y = y0[j]; y < y1[j]; ++y) {
                    const int ey0 = r.bin_y[y], ey1 = r.bin_y[y + 1];
                    for (int x = x0[j]; x < x1[j]; ++x) add(x, y, r.bin_x[x], ey0, r.bin_x[x + 1], ey1);
                }
            }
        }
        __syncthreads();
        if (dense) {
            for (int i = threadIdx.x; i < nd; i += kBlock) {
                const V a = d_inst[i], b = d_fill[i];
                if (!(a | b)) continue;
                const int yy = i / BW;
                const unsigned kb = (unsigned)(by0 + yy) * (unsigned)r.cnt_x + (unsigned)(bx0 + i - yy * BW);
                if (a) atomicAdd(r.inst + kb, (unsigned long long)a);
                if (b) atomicAdd(r.filler + kb, (unsigned long long)b);
                adds += (a != 0) + (b != 0);
            }
        } else {
            for (int i = threadIdx.x; i < kSlots; i += kBlock) {
                const unsigned kb = t_key[i];
                if (kb == kEmpty) continue;
                const V a = t_inst[i], b = t_fill[i];
                if (a) atomicAdd(r.inst + kb, (unsigned long long)a);
                if (b) atomicAdd(r.filler + kb, (unsigned long long)b);
                adds += (a != 0) + (b != 0);
            }
        }
        if (adds) atomicAdd(&s_adds, adds);
        __syncthreads();
    }
    if (threadIdx.x == 0 && s_adds && pa.global_adds) atomicAdd(pa.global_adds, (unsigned long long)s_adds);
}

template <int MODE, typename V, int CH>
__global__ void __launch_bounds__(kBlock) k_place_scatter(RegionDev r, CellsDev c, PlaceArgs pa) {
    place_scatter_body<MODE, V, CH>(r, c, pa, blockIdx.x, gridDim.x);
}

// The finishing reduction of a kernel's FP64 block partials into *out, in a fixed order (one
// block): quantities 0-1 over nb_bins blocks (the overflow sums), 2-5 over nb_cells (the gradient
// and step-length sums); the HPWL counter copied when given.
__device__ void finish_sums(const double* partials, int nb_bins, int nb_cells, Sums* out, const unsigned long long* hpwl, double* sh) {
    double v[6];
    for (int q = 0; q < 6; ++q) {
        int cnt = q < 2 ? nb_bins : nb_cells;
        double s = 0.0;
        for (int i = threadIdx.x; i < cnt; i += blockDim.x) s += partials[q * kPartials + i];
        v[q] = block_sum(s, sh);
    }
    if (threadIdx.x == 0) {
        if (nb_bins > 0) {
            out->overflow = v[0];
            out->overflow_unscaled = v[1];
        }
        if (nb_cells > 0) {
            out->wl_abs = v[2];
            out->den_abs = v[3];
            out->coord_sq = v[4];
            out->grad_sq = v[5];
        }
        if (hpwl) out->hpwl = *hpwl;
    }
}

// LDS pointers typed as such, so that every access compiles to ds_read / ds_write (a generic
// pointer into LDS makes flat accesses, which wait like global loads).
#define GPL_LDS __attribute__((address_space(3)))
typedef float vf2 __attribute__((ext_vector_type(2)));
typedef GPL_LDS vf2 lds_float2;
typedef GPL_LDS float lds_float;
__device__ __forceinline__ float2 ld2(const lds_float2* p) {
    const vf2 v = *p;
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ float2 ld2(const float2* p) { return *p; }
__device__ __forceinline__ void st2(lds_float2* p, float2 v) {
    vf2 t;
    t.x = v.x;
    t.y = v.y;
    *p = t;
}

// The bins pass's layout, shared with the solve's first row pass (which runs it fused): a block of
// kBlock threads takes rows_of_block consecutive bin rows, row_threads threads per row (the row
// transform's L/8), each thread the elements x = tid + i * row_threads; its FP64 overflow terms are
// summed in that order, then over the block in a fixed tree, one partial per block.  The same
// layout in both kernels gives the same partials bit for bit.
__host__ __device__ constexpr int row_threads(int L) { return (L >> 1) >= 4 ? (L >> 1) >> 2 : 1; }
__host__ __device__ constexpr int rows_of_block(int L) { return kBlock / row_threads(L) > 0 ? kBlock / row_threads(L) : 1; }

// updateBinsGCellDensityArea's per-bin part (nesterov.rs:2252-2262) over one thread's elements of
// bin row y: density (to dens, and with ROW to the LDS row), the overflow terms added to o and ou,
// the accumulators cleared when asked.
template <bool ROW>
__device__ __forceinline__ void bins_row_terms(unsigned long long* inst_a, unsigned long long* fill_a, const float2* npf_a, const int32_t* bin_x,
                                               const int32_t* bin_y, float td, float* dens, lds_float* rowv, int cnt_x, int y, int tid,
                                               int T, bool clear, double& o, double& ou) {
    const size_t base = (size_t)y * (size_t)cnt_x;
    const int ey = bin_y[y + 1] - bin_y[y];
    // eight elements at a time, every load issued before the first store (the stores could alias
    // the next loads, so a fused loop would wait out each load in turn)
    constexpr int kPer = 8;
    for (int x0 = tid; x0 < cnt_x; x0 += kPer * T) {
        long long iv[kPer], fv[kPer];
        float2 nv[kPer];
        int bw[kPer];
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const int x = x0 + i * T;
            if (x < cnt_x) {
                const size_t k = base + (size_t)x;
                iv[i] = (long long)inst_a[k];
                fv[i] = (long long)fill_a[k];
                nv[i] = npf_a[k];  // the two int64 constants, converted once on the host
                bw[i] = bin_x[x + 1] - bin_x[x];
            }
        }
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const int x = x0 + i * T;
            if (x >= cnt_x) break;
            const size_t k = base + (size_t)x;
            const long long area = (long long)bw[i] * (long long)ey;
            const float scaled = (float)area * td;
            const float d = ((float)iv[i] + (float)fv[i] + nv[i].x) / scaled;
            dens[k] = d;
            if (ROW) rowv[x] = d;
            float t = rmax((float)iv[i] + nv[i].x - scaled, 0.0f);
            float tu = rmax((float)iv[i] + nv[i].y - scaled, 0.0f);
            if (!(t >= 0.0f)) t = 0.0f;
            if (!(tu >= 0.0f)) tu = 0.0f;
            o += (double)t;
            ou += (double)tu;
            if (clear) {
                inst_a[k] = 0ull;
                fill_a[k] = 0ull;
            }
        }
    }
}

// The bins pass alone: density, the overflow partials, the accumulators cleared for the next
// scatter (unless `clear` is 0).
__global__ void __launch_bounds__(kBlock) k_bin_density(RegionDev r, double* partials, int clear) {
    if (r.ctl && (r.ctl->hot.flags & 1)) return;
    __builtin_amdgcn_s_setprio(2);  // on the trial's critical chain (as the scatter)
    __shared__ double sh[kBlock / 64];
    const int T = row_threads(r.cnt_x), rpb = rows_of_block(r.cnt_x);
    const int sub = (int)threadIdx.x / T, tid = (int)threadIdx.x - sub * T;
    const int y = (int)blockIdx.x * rpb + sub;
    double o = 0.0, ou = 0.0;
    if (sub < rpb && y < r.cnt_y)
        bins_row_terms<false>(r.inst, r.filler, r.non_place_f, r.bin_x, r.bin_y, r.target_density, r.density, nullptr, r.cnt_x, y, tid, T,
                       clear != 0, o, ou);
    double so = block_sum(o, sh);
    double sou = block_sum(ou, sh);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = so;
        partials[kPartials + blockIdx.x] = sou;
    }
}

// ---------------------------------------------------------------------------------------------
// The Poisson solve: 1-D cosine / sine transforms along rows, L/8 threads per row and the row in
// LDS; an L/2-point complex Stockham FFT (radix 4, one radix-2 pass when log2(L/2) is odd) with
// Makhoul's mapping between the real transform and the FFT.  Conventions are Ooura's
// (ooura.rs:158-255):
//   DCT-II   C[k] = sum_j a[j] cos(pi (j+1/2) k / L)
//   DCT-III  C[k] = sum_j a[j] cos(pi j (k+1/2) / L)
//   DST-III  S[k] = sum_{j=1..L} A[j] sin(pi j (k+1/2) / L), A[L] in a[0]
// ---------------------------------------------------------------------------------------------

enum RowMode {
    ROW_DCT2 = 0,
    ROW_DCT2_SCALE = 1,
    ROW_POISSON = 2,
    ROW_INV_FIELDS = 3,
    ROW_INV_PHI = 4,
    ROW_SPECTRUM_POISSON = 6,
    ROW_BINS_DCT2 = 7
};

// The modes that run two inverse transforms side by side (a second pair of FFT buffers per row).
__host__ __device__ constexpr bool row_mode_dual(int m) { return m == ROW_POISSON || m == ROW_INV_FIELDS || m == ROW_SPECTRUM_POISSON; }

__device__ __forceinline__ float2 cmul(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }


// Complex FFTs of N points (N = 2^logN, or 3 * 2^logN with r3), NB of them side by side (each held
// in LDS buffer x[b]; the NB transforms share the twiddles and the barriers), by T = max(N/4, 1)
// threads of a row group (tid < T), Stockham autosort: radix-2 first when logN is odd, then radix-4
// stages, then with r3 one radix-3 stage (its sub-transforms of length 2^logN, twiddles w^k and
// w^2k); natural order in and out.  tw[m] = e^{-2 pi i m / N}, m < N (TP: an LDS or a global
// pointer); dir -1 forward, +1 inverse (unnormalised).  LOGN > 0 fixes N = 2^LOGN at compile time
// (the stages unroll, the index arithmetic folds); 0 takes N, logN and r3 as given.  Every thread
// of the block calls it (block barriers inside); on return x[b] is the buffer that holds result b.
template <int LOGN, int NB, typename TP>
__device__ void fft_stockham_n(lds_float2* (&x)[NB], lds_float2* (&y)[NB], int N_, int logN_, int T_, int tid, TP tw, int dir,
                               int r3 = 0) {
    const int logN = LOGN > 0 ? LOGN : logN_;
    const int N = LOGN > 0 ? (1 << LOGN) : N_;
    const int T = LOGN > 0 ? (LOGN >= 2 ? (1 << (LOGN - 2)) : 1) : T_;
    int lNs = 0;  // log2 of the current sub-transform length
    auto swap_all = [&]() {
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            lds_float2* t = x[b];
            x[b] = y[b];
            y[b] = t;
        }
    };
    auto radix2 = [&](int j, int half) {
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            float2 v0 = ld2(x[b] + j), v1 = ld2(x[b] + j + half);
            st2(y[b] + 2 * j, make_float2(v0.x + v1.x, v0.y + v1.y));
            st2(y[b] + 2 * j + 1, make_float2(v0.x - v1.x, v0.y - v1.y));
        }
    };
    __syncthreads();
    if (logN & 1) {
        const int half = N >> 1;
        // LOGN > 0: half / T = 2 passes (1 when N = 2); LOGN == 0: the loop after this one
#pragma unroll
        for (int i = 0; i < (LOGN >= 2 ? 2 : 1); ++i) {
            const int j = tid + i * T;
            if (LOGN > 0 && j < half) radix2(j, half);
        }
        if (LOGN == 0)
            for (int j = tid; j < half; j += T) radix2(j, half);
        lNs = 1;
        swap_all();
        __syncthreads();
    }
    const int q = N >> 2;
#pragma unroll
    for (int st = 0; st < (LOGN > 0 ? LOGN / 2 : 16); ++st) {
        if (LOGN == 0 && lNs >= logN) break;
        const int Ns = 1 << lNs;
        if (tid < q) {
            const int j = tid;
            const int k = j & (Ns - 1);
            float2 v[NB][4];
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                v[b][0] = ld2(x[b] + j);
                v[b][1] = ld2(x[b] + j + q);
                v[b][2] = ld2(x[b] + j + 2 * q);
                v[b][3] = ld2(x[b] + j + 3 * q);
            }
            if (Ns > 1) {
                const int step = N >> (lNs + 2);  // twiddle index stride
                float2 w1 = ld2(tw + k * step), w2 = ld2(tw + 2 * k * step), w3 = ld2(tw + 3 * k * step);
                if (dir > 0) {
                    w1 = cconj(w1);
                    w2 = cconj(w2);
                    w3 = cconj(w3);
                }
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    v[b][1] = cmul(v[b][1], w1);
                    v[b][2] = cmul(v[b][2], w2);
                    v[b][3] = cmul(v[b][3], w3);
                }
            }
            const int d = ((j >> lNs) << (lNs + 2)) + k;
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const float2 v0 = v[b][0], v1 = v[b][1], v2 = v[b][2], v3 = v[b][3];
                float2 t0 = make_float2(v0.x + v2.x, v0.y + v2.y), t1 = make_float2(v0.x - v2.x, v0.y - v2.y);
                float2 t2 = make_float2(v1.x + v3.x, v1.y + v3.y), t3 = make_float2(v1.x - v3.x, v1.y - v3.y);
                // forward: X1 = t1 - i t3, X3 = t1 + i t3; inverse the other way
                float2 it3 = dir < 0 ? make_float2(t3.y, -t3.x) : make_float2(-t3.y, t3.x);  // -i t3 or +i t3
                st2(y[b] + d, make_float2(t0.x + t2.x, t0.y + t2.y));
                st2(y[b] + d + Ns, make_float2(t1.x + it3.x, t1.y + it3.y));
                st2(y[b] + d + 2 * Ns, make_float2(t0.x - t2.x, t0.y - t2.y));
                st2(y[b] + d + 3 * Ns, make_float2(t1.x - it3.x, t1.y - it3.y));
            }
        }
        lNs += 2;
        swap_all();
        __syncthreads();
    }
    if (LOGN == 0 && r3) {
        // the radix-3 stage: X0 = a + b + c, X1,2 = a - (b + c)/2 -/+ i (sqrt(3)/2)(b - c) forward
        // (the signs swapped inverse), b and c twiddled by w^k, w^2k
        const int third = N / 3;
        const float s3 = 0.86602540378443864676f;
        for (int j = tid; j < third; j += T) {
            float2 w1 = ld2(tw + j), w2 = ld2(tw + 2 * j);
            if (dir > 0) {
                w1 = cconj(w1);
                w2 = cconj(w2);
            }
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const float2 a0 = ld2(x[b] + j);
                const float2 b0 = cmul(ld2(x[b] + j + third), w1), c0 = cmul(ld2(x[b] + j + 2 * third), w2);
                const float2 t1 = make_float2(b0.x + c0.x, b0.y + c0.y);
                const float2 t2 = make_float2(a0.x - 0.5f * t1.x, a0.y - 0.5f * t1.y);
                const float2 t3 = make_float2(s3 * (b0.x - c0.x), s3 * (b0.y - c0.y));
                const float2 m = make_float2(t2.x + t3.y, t2.y - t3.x);  // t2 - i t3
                const float2 p = make_float2(t2.x - t3.y, t2.y + t3.x);  // t2 + i t3
                st2(y[b] + j, make_float2(a0.x + t1.x, a0.y + t1.y));
                st2(y[b] + j + third, dir < 0 ? m : p);
                st2(y[b] + j + 2 * third, dir < 0 ? p : m);
            }
        }
        swap_all();
        __syncthreads();
    }
}

template <int LOGN, typename TP>
__device__ lds_float2* fft_stockham(lds_float2* x, lds_float2* y, int N_, int logN_, int T_, int tid, TP tw, int dir, int r3 = 0) {
    lds_float2* xs[1] = {x};
    lds_float2* ys[1] = {y};
    fft_stockham_n<LOGN, 1>(xs, ys, N_, logN_, T_, tid, tw, dir, r3);
    return xs[0];
}

// A row's transform context.  LOGL > 0 fixes the row length L = 2^LOGL (and with it h = L/2, the
// T = h/4 threads per row, the FFT's stages) at compile time; 0 takes them from the fields (h =
// 2^logh, or 3 * 2^logh with r3).
template <typename TP, int LOGL>
struct RowCtx {
    int L_, h_, logh_, T_, r3_, tid;
    TP tw;  // e^{-2 pi i m / h}, m < h
    TP th;  // e^{-2 pi i k / L}, k <= h
    TP dw;  // (cos, -sin)(pi k / 2L), k < L
    lds_float2* A;
    lds_float2* B;
    __device__ __forceinline__ int L() const { return LOGL > 0 ? (1 << LOGL) : L_; }
    __device__ __forceinline__ int h() const { return LOGL > 0 ? (1 << (LOGL - 1)) : h_; }
    __device__ __forceinline__ int logh() const { return LOGL > 0 ? LOGL - 1 : logh_; }
    __device__ __forceinline__ int T() const { return LOGL > 0 ? (LOGL >= 3 ? (1 << (LOGL - 3)) : 1) : T_; }
    __device__ __forceinline__ int r3() const { return LOGL > 0 ? 0 : r3_; }
};

// DCT-II of the real row given by src(j), j < L (Makhoul: v = (x[0], x[2], ..., x[3], x[1]), its
// DFT through an L/2-point complex FFT of the pairs, C[k] = Re(e^{-i pi k / 2L} V[k])); calls
// out(k, C[k]).  (h is a multiple of T: every loop below runs h / T or L / T times per thread.)
template <typename TP, int LOGL, typename SRC, typename OUT>
__device__ void dct2_row(const RowCtx<TP, LOGL>& R, bool active, SRC src, OUT out) {
    const int L = R.L(), h = R.h(), T = R.T();
    if (active) {
#pragma unroll
        for (int i = 0; i < h / T; ++i) {
            const int m = R.tid + i * T;
            int q0 = 2 * m, q1 = 2 * m + 1;
            float a = q0 < h ? src(2 * q0) : src(2 * (L - 1 - q0) + 1);
            float b = q1 < h ? src(2 * q1) : src(2 * (L - 1 - q1) + 1);
            st2(R.A + m, make_float2(a, b));
        }
    }
    lds_float2* Z = fft_stockham<(LOGL > 0 ? LOGL - 1 : 0)>(R.A, R.B, h, R.logh(), T, R.tid, R.tw, -1, R.r3());
    if (active) {
        auto post = [&](int k) {
            float2 zk = ld2(Z + (k == h ? 0 : k)), zc = cconj(ld2(Z + (k == 0 ? 0 : h - k)));
            float2 E = make_float2((zk.x + zc.x) * 0.5f, (zk.y + zc.y) * 0.5f);
            float2 Op = make_float2((zk.x - zc.x) * 0.5f, (zk.y - zc.y) * 0.5f);
            float2 O = make_float2(Op.y, -Op.x);  // -i Op
            float2 wO = cmul(ld2(R.th + k), O);
            float2 V = make_float2(E.x + wO.x, E.y + wO.y);
            const float2 d0 = ld2(R.dw + k);
            out(k, d0.x * V.x - d0.y * V.y);
            if (k > 0 && k < h) {
                const float2 d1 = ld2(R.dw + (L - k));
                out(L - k, d1.x * V.x + d1.y * V.y);
            }
        };
#pragma unroll
        for (int i = 0; i < h / T; ++i) post(R.tid + i * T);
        if (R.tid == 0) post(h);
    }
}

// DCT-III (sine: DST-III) of the row given by coef(j), j < L (Ooura's conventions); calls
// out(j, y[j]).  V[q] = e^{i pi q / 2L} (c[q] - i c[L-q]) with c[0] = 2 a[0], c[L] = 0 (the sine:
// c from b[0] = a[0], b[j] = a[L-j]); the real inverse DFT through an L/2-point complex IFFT of
// Z[k] = (V[k] + conj V[h-k]) + i e^{2 pi i k / L} (V[k] - conj V[h-k]); y = v / 2 unshuffled,
// odd outputs negated for the sine.
template <typename TP, int LOGL, typename F>
__device__ __forceinline__ void dct3_pre(const RowCtx<TP, LOGL>& R, lds_float2* dst, F coef, bool sine) {
    const int L = R.L(), h = R.h(), T = R.T();
    auto V = [&](int q) -> float2 {
        float ck, cl;
        if (!sine) {
            ck = q == 0 ? 2.0f * coef(0) : coef(q);
            cl = q == 0 ? 0.0f : coef(L - q);
        } else {
            ck = q == 0 ? 2.0f * coef(0) : coef(L - q);
            cl = q == 0 ? 0.0f : coef(q);
        }
        const float2 dq = ld2(R.dw + q);
        float cs = dq.x, sn = -dq.y;
        return make_float2(cs * ck + sn * cl, sn * ck - cs * cl);
    };
#pragma unroll
    for (int i = 0; i < h / T; ++i) {
        const int k = R.tid + i * T;
        float2 vk = V(k), vc = cconj(V(h - k));
        float2 E = make_float2(vk.x + vc.x, vk.y + vc.y), D = make_float2(vk.x - vc.x, vk.y - vc.y);
        float2 wD = cmul(cconj(ld2(R.th + k)), D);
        st2(dst + k, make_float2(E.x - wD.y, E.y + wD.x));
    }
}
template <typename TP, int LOGL, typename OUT>
__device__ __forceinline__ void dct3_post(const RowCtx<TP, LOGL>& R, const lds_float2* z, bool sine, OUT out) {
    const int L = R.L(), T = R.T();
#pragma unroll
    for (int i = 0; i < L / T; ++i) {
        const int j = R.tid + i * T;
        int m = (j & 1) ? (L - 1 - (j >> 1)) : (j >> 1);
        float2 zz = ld2(z + (m >> 1));
        float v = ((m & 1) ? zz.y : zz.x) * 0.5f;
        if (sine && (j & 1)) v = -v;
        out(j, v);
    }
}

// DCT-III (sine: DST-III) of the row given by coef(j), j < L (Ooura's conventions); calls
// out(j, y[j]).  V[q] = e^{i pi q / 2L} (c[q] - i c[L-q]) with c[0] = 2 a[0], c[L] = 0 (the sine:
// c from b[0] = a[0], b[j] = a[L-j]); the real inverse DFT through an L/2-point complex IFFT of
// Z[k] = (V[k] + conj V[h-k]) + i e^{2 pi i k / L} (V[k] - conj V[h-k]); y = v / 2 unshuffled,
// odd outputs negated for the sine.
template <typename TP, int LOGL, typename F, typename OUT>
__device__ void dct3_row(const RowCtx<TP, LOGL>& R, bool active, F coef, bool sine, OUT out) {
    if (active) dct3_pre(R, R.A, coef, sine);
    lds_float2* z = fft_stockham<(LOGL > 0 ? LOGL - 1 : 0)>(R.A, R.B, R.h(), R.logh(), R.T(), R.tid, R.tw, +1, R.r3());
    if (active) dct3_post(R, z, sine, out);
    __syncthreads();
}

// Two DCT-III / DST-III of one row side by side (the second in the buffers A1, B1): one FFT stage
// loop, one barrier per stage for both.  Each output is the same as dct3_row's.
template <typename TP, int LOGL, typename F0, typename OUT0, typename F1, typename OUT1>
__device__ void dct3_row2(const RowCtx<TP, LOGL>& R, lds_float2* A1, lds_float2* B1, bool active, F0 coef0, bool sine0, OUT0 out0,
                          F1 coef1, bool sine1, OUT1 out1) {
    if (active) {
        dct3_pre(R, R.A, coef0, sine0);
        dct3_pre(R, A1, coef1, sine1);
    }
    lds_float2* x[2] = {R.A, A1};
    lds_float2* y[2] = {R.B, B1};
    fft_stockham_n<(LOGL > 0 ? LOGL - 1 : 0), 2>(x, y, R.h(), R.logh(), R.T(), R.tid, R.tw, +1, R.r3());
    if (active) {
        dct3_post(R, x[0], sine0, out0);
        dct3_post(R, x[1], sine1, out1);
    }
    __syncthreads();
}

struct RowArgs {
    const float* in0;
    const float* in1;
    float* out0;
    float* out1;
    float* out2;
    int L, logh, nrows, T;
    int r3;              // the half length h = 3 * 2^logh (else 2^logh)
    const float2* tw;
    const float2* th;
    const float2* dw;
    const float* wrow;   // wavenumber of the row index
    const float* wcol;   // wavenumber of the element index
    double scale;
    int want_phi;
    const int* skip;     // when set and *skip: the launch does nothing (a device-driven trial's halted region)
    int transpose_out;   // outputs written transposed: out[k * nrows + row]
    int interleave;      // ROW_SPECTRUM_POISSON: the two fields as pairs (Ex', Ey') in out0, [k][row][2];
                         // ROW_INV_FIELDS: in0 holds such pairs, [row][v][2]
    float2* pairs;       // ROW_INV_FIELDS: the two outputs as (out0, out1) pairs, [row][j] (out0/out1 unused)
    // ROW_BINS_DCT2: the bins pass's inputs and outputs (k_bin_density's), the rows being bin rows
    unsigned long long* inst;
    unsigned long long* filler;
    const float2* npf;
    const int32_t* bin_x;
    const int32_t* bin_y;
    float target;
    float* density;
    double* partials;
};

// Row transforms, L/8 threads per row, 256 / (L/8) rows per block (rows of one XCD contiguous,
// so that the blocks writing the same lines of a transposed output share an L2).  Modes:
//  ROW_DCT2        out0[row] = DCT-II(in0[row])
//  ROW_DCT2_SCALE  the same, then the halving of row/column 0 and x 4/(nx ny) (nesterov.rs:415-424)
//  ROW_POISSON     in0 = scaled spectrum rows A[u][v] (u = row): Ex = p wx[u], Ey = p wy[v],
//                  p = A / (wx^2 + wy^2) (nesterov.rs:425-440); out0 = DCT-III(Ex), out1 =
//                  DST-III(Ey), out2 = DCT-III(p) when want_phi
//  ROW_INV_FIELDS  out0 = DST-III(in0), out1 = DCT-III(in1)
//  ROW_INV_PHI     out0 = DCT-III(in0)
//  ROW_SPECTRUM_POISSON  ROW_DCT2_SCALE kept in LDS, then ROW_POISSON on it
//  ROW_BINS_DCT2   the bins pass of the row (k_bin_density's per-bin part and partials, in its
//                  layout: L a power of two, so that a block is kBlock threads) into LDS, then
//                  ROW_DCT2 of it
// With transpose_out the outputs go out transposed, so that the next pass reads rows.
// TL: the twiddle tables (and the element wavenumbers) are first copied to LDS (the launcher
// picks it when they fit beside the rows' buffers); else they are read from global memory.
// LDS: per row the FFT's two buffers (L float2), [the modes with two inverse transforms side by
// side: the second transform's two buffers, L float2 per row], [ROW_SPECTRUM_POISSON: the spectrum, L floats],
// then [TL: tw h, th h + 1, dw L float2, the element wavenumbers L floats].
template <int MODE, bool TL, int LOGL>
__device__ __forceinline__ void rows_body(RowArgs a, const unsigned bid_, const unsigned nblk_) {
    if (a.skip && (*a.skip & 1)) return;
    __builtin_amdgcn_s_setprio(2);  // on the trial's critical chain (as the scatter)
    extern __shared__ float2 lds2[];
    lds_float2* const lbase = (lds_float2*)lds2;
    using TP = typename std::conditional<TL, const lds_float2*, const float2*>::type;
    RowCtx<TP, LOGL> R;
    R.L_ = a.L;
    R.h_ = a.L >> 1;
    R.logh_ = a.logh;
    R.T_ = a.T;
    R.r3_ = a.r3;
    const int T = R.T(), L = R.L(), h = R.h();
    const int rpb = LOGL > 0 ? 256 / T : (int)blockDim.x / T;
    const int sub = threadIdx.x / T;
    R.tid = threadIdx.x - sub * T;
    R.A = lbase + (size_t)sub * L;
    R.B = R.A + h;
    constexpr bool dual = row_mode_dual(MODE);
    lds_float2* const dbase = lbase + (size_t)rpb * L;  // [dual] the second transform's buffers
    lds_float2* const A1 = dbase + (size_t)sub * L;
    lds_float2* const B1 = A1 + h;
    lds_float2* const after = dual ? dbase + (size_t)rpb * L : dbase;
    lds_float* extra = (lds_float*)after + (size_t)sub * L;
    const bool spectrum = MODE == ROW_SPECTRUM_POISSON;
    const bool wav = MODE == ROW_POISSON || MODE == ROW_SPECTRUM_POISSON;
    lds_float2* tab = after + (spectrum ? (size_t)rpb * L / 2 : 0);
    lds_float* wcol_l = (lds_float*)(tab + 2 * h + 1 + L);
    if constexpr (TL) {
        for (int i = threadIdx.x; i < h; i += blockDim.x) st2(tab + i, a.tw[i]);
        for (int i = threadIdx.x; i <= h; i += blockDim.x) st2(tab + h + i, a.th[i]);
        for (int i = threadIdx.x; i < L; i += blockDim.x) st2(tab + 2 * h + 1 + i, a.dw[i]);
        if (wav)
            for (int i = threadIdx.x; i < L; i += blockDim.x) wcol_l[i] = a.wcol[i];
        R.tw = tab;
        R.th = tab + h;
        R.dw = tab + 2 * h + 1;
        __syncthreads();
    } else {
        R.tw = a.tw;
        R.th = a.th;
        R.dw = a.dw;
    }
    auto wcol = [&](int v) -> float {
        if constexpr (TL) return wcol_l[v];
        else return a.wcol[v];
    };
    const int lb = nblk_ % kXcds == 0 ? xcd_block(bid_, nblk_) : (int)bid_;
    const int row = lb * rpb + sub;
    const bool active = row < a.nrows;
    const size_t off = (size_t)(active ? row : 0) * L;
    const int nrows = a.nrows;
    const bool tr = a.transpose_out != 0;
    auto put = [&](float* o, int k, float v) {
        if (tr) o[(size_t)k * nrows + row] = v;
        else o[off + k] = v;
    };
    // Transposed outputs when every row of every block is present: each row's outputs staged in
    // its FFT buffer that the transform left free (the result lies in B after an odd number of
    // Stockham swaps), then the block's rpb consecutive rows written together, rpb floats per
    // output column (a 4-byte store per element into a different line otherwise).
    const bool group = tr && rpb >= 2 && (rpb & 1) == 0 && nrows % rpb == 0;
    const int lh = R.logh();
    const bool res_b = (((lh & 1) + (lh >> 1) + R.r3()) & 1) != 0;
    auto stage_of = [&](bool second, int s) -> lds_float* {
        lds_float2* base = (second ? dbase : lbase) + (size_t)s * L;
        return (lds_float*)(res_b ? base : base + h);
    };
    lds_float* const st0 = stage_of(false, sub);
    lds_float* const st1 = stage_of(true, sub);
    auto emit = [&](float* o, lds_float* st, int k, float v) {
        if (group) st[k] = v;
        else put(o, k, v);
    };
    auto flush = [&](float* o, bool second) {
        if (!group) return;
        __syncthreads();
        const int row0 = lb * rpb, np = rpb / 2;
        for (int i = threadIdx.x; i < L * np; i += blockDim.x) {
            const int k = i / np, pr = i - k * np;
            const float2 v = make_float2(stage_of(second, 2 * pr)[k], stage_of(second, 2 * pr + 1)[k]);
            *reinterpret_cast<float2*>(o + (size_t)k * nrows + row0 + 2 * pr) = v;
        }
        __syncthreads();
    };
    if (MODE == ROW_BINS_DCT2) {
        // the row's densities into the FFT buffer that dct2_row fills last (B: its first loop
        // writes A, and the transform's first barrier comes before any write into B)
        __shared__ double sh[kBlock / 64];
        lds_float* rowv = (lds_float*)R.B;
        double o = 0.0, ou = 0.0;
        if (active)
            bins_row_terms<true>(a.inst, a.filler, a.npf, a.bin_x, a.bin_y, a.target, a.density, rowv, L, row, R.tid, T, true, o, ou);
        const double so = block_sum(o, sh);
        const double sou = block_sum(ou, sh);
        if (threadIdx.x == 0) {
            a.partials[lb] = so;
            a.partials[kPartials + lb] = sou;
        }
        __syncthreads();
        dct2_row(R, active, [&](int j) { return rowv[j]; }, [&](int k, float v) { emit(a.out0, st0, k, v); });
        flush(a.out0, false);
    } else if (MODE == ROW_DCT2 || MODE == ROW_DCT2_SCALE) {
        const double scale = a.scale;
        const float* src = a.in0 + off;
        dct2_row(R, active, [&](int j) { return src[j]; }, [&](int k, float v) {
            if (MODE == ROW_DCT2_SCALE) {
                if ((row >= 1 && k == 0) || (row == 0 && k >= 1)) v = (float)((double)v * 0.5);
                v = (float)((double)v * scale);
            }
            emit(a.out0, st0, k, v);
        });
        flush(a.out0, false);
    } else if (MODE == ROW_SPECTRUM_POISSON) {
        // the spectrum row, halved and scaled, then the potential p = A / (wx^2 + wy^2) of each
        // element (DC 0) in place: both field transforms read it
        const double scale = a.scale;
        const float* Ag = a.in0 + off;
        dct2_row(R, active, [&](int j) { return Ag[j]; }, [&](int k, float v) {
            if ((row >= 1 && k == 0) || (row == 0 && k >= 1)) v = (float)((double)v * 0.5);
            v = (float)((double)v * scale);
            extra[k] = v;
        });
        __syncthreads();
        const float wu = active ? a.wrow[row] : 0.0f, wu2 = wu * wu;
        if (active) {
#pragma unroll
            for (int i = 0; i < L / T; ++i) {
                const int v = R.tid + i * T;
                const float wv = wcol(v);
                extra[v] = (row == 0 && v == 0) ? 0.0f : extra[v] / (wu2 + wv * wv);
            }
        }
        __syncthreads();
        const bool il = a.interleave != 0;
        dct3_row2(
            R, A1, B1, active, [&](int v) { return extra[v] * wu; }, false,
            [&](int j, float y) {
                if (group) st0[j] = y;
                else if (il) a.out0[((size_t)j * nrows + row) * 2] = y;
                else put(a.out0, j, y);
            },
            [&](int v) { return extra[v] * wcol(v); }, true,
            [&](int j, float y) {
                if (group) st1[j] = y;
                else if (il) a.out0[((size_t)j * nrows + row) * 2 + 1] = y;
                else put(a.out1, j, y);
            });
        if (a.interleave) {
            if (group) {
                // the block's rows' (Ex', Ey') pairs, 2 rpb floats per output column
                __syncthreads();
                const int row0 = lb * rpb, np = rpb / 2;
                for (int i = threadIdx.x; i < L * np; i += blockDim.x) {
                    const int k = i / np, pr = i - k * np;
                    const float4 v = make_float4(stage_of(false, 2 * pr)[k], stage_of(true, 2 * pr)[k], stage_of(false, 2 * pr + 1)[k],
                                                 stage_of(true, 2 * pr + 1)[k]);
                    *reinterpret_cast<float4*>(a.out0 + ((size_t)k * nrows + row0 + 2 * pr) * 2) = v;
                }
                __syncthreads();
            }
        } else {
            flush(a.out0, false);
            flush(a.out1, true);
        }
        if (a.want_phi) {
            dct3_row(R, active, [&](int v) { return extra[v]; }, false, [&](int j, float y) { emit(a.out2, st0, j, y); });
            flush(a.out2, false);
        }
    } else if (MODE == ROW_POISSON) {
        const float* A = a.in0 + off;
        const float wu = active ? a.wrow[row] : 0.0f, wu2 = wu * wu;
        auto pot = [&](int v) -> float {
            if (row == 0 && v == 0) return 0.0f;
            float wv = wcol(v);
            return A[v] / (wu2 + wv * wv);
        };
        dct3_row2(
            R, A1, B1, active, [&](int v) { return pot(v) * wu; }, false, [&](int j, float y) { emit(a.out0, st0, j, y); },
            [&](int v) { return pot(v) * wcol(v); }, true, [&](int j, float y) { emit(a.out1, st1, j, y); });
        flush(a.out0, false);
        flush(a.out1, true);
        if (a.want_phi) {
            dct3_row(R, active, pot, false, [&](int j, float y) { emit(a.out2, st0, j, y); });
            flush(a.out2, false);
        }
    } else if (MODE == ROW_INV_FIELDS) {
        const int il = a.interleave ? 2 : 1;
        const float* X = il == 2 ? a.in0 + 2 * off : a.in0 + off;
        const float* Y = il == 2 ? a.in0 + 2 * off + 1 : a.in1 + off;
        if (a.pairs) {
            // both outputs staged in the freed FFT buffers, then the row written as (x, y) pairs
            dct3_row2(
                R, A1, B1, active, [&](int v) { return X[il * v]; }, true, [&](int j, float y) { st0[j] = y; },
                [&](int v) { return Y[il * v]; }, false, [&](int j, float y) { st1[j] = y; });
            if (active)
                for (int j = R.tid; j < L; j += T) a.pairs[off + j] = make_float2(st0[j], st1[j]);
        } else {
            dct3_row2(
                R, A1, B1, active, [&](int v) { return X[il * v]; }, true, [&](int j, float y) { emit(a.out0, st0, j, y); },
                [&](int v) { return Y[il * v]; }, false, [&](int j, float y) { emit(a.out1, st1, j, y); });
            flush(a.out0, false);
            flush(a.out1, true);
        }
    } else if (MODE == ROW_INV_PHI) {
        const float* X = a.in0 + off;
        dct3_row(R, active, [&](int v) { return X[v]; }, false, [&](int j, float y) { emit(a.out0, st0, j, y); });
        flush(a.out0, false);
    }
}

template <int MODE, bool TL, int LOGL>
__global__ void __launch_bounds__(256) k_rows(RowArgs a) {
    rows_body<MODE, TL, LOGL>(a, blockIdx.x, gridDim.x);
}

// Tiled transpose: in [R][C] -> out [C][R].
__global__ void k_transpose(const float* in, float* out, int R, int C) {
    __shared__ float tile[32][33];
    int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    for (int j = threadIdx.y; j < 32; j += blockDim.y) {
        int r = by + j, c = bx + threadIdx.x;
        if (r < R && c < C) tile[j][threadIdx.x] = in[(size_t)r * C + c];
    }
    __syncthreads();
    for (int j = threadIdx.y; j < 32; j += blockDim.y) {
        int c = bx + j, r = by + threadIdx.x;
        if (r < R && c < C) out[(size_t)c * R + r] = tile[threadIdx.x][j];
    }
}

// ---------------------------------------------------------------------------------------------
// WA wirelength (nesterov.rs:2325-2430)
// ---------------------------------------------------------------------------------------------

// A net pin's position (nesterov.rs:2150, 2268-2276) from its cell index, the cell's centre and
// the pin's offset: its cell's centre plus the offset; a pin on no movable cell (port, fixed
// instance) holds its fixed position in the offset (the centre, read from cell 0, is then
// unused); a cell not yet centred leaves its pins where init put them.  The centre is read
// unconditionally so that the callers can issue every pin's reads before waiting for any.
__device__ __forceinline__ int2 pin_at(const NetsDev& n, int k, int cell, int2 m, int2 o) {
    if (cell < 0) return o;
    if (m.x == INT_MIN) return n.pxy0[k];
    return make_int2(m.x + o.x, m.y + o.y);
}
__device__ __forceinline__ int2 pin_pos(const NetsDev& n, const CellsDev& c, int k) {
    const int cell = n.pcell[k];
    const int2 o = n.poff[k];
    return pin_at(n, k, cell, c.ctr[max(cell, 0)], o);
}

// Pins s + i * stride (i < R, i * stride < deg) of a net: cell indices, offsets and, with GRAD,
// cell-pin indices read in one batch, then every centre in a second, so that a net waits for two
// memory round trips instead of two or three per pin.
template <int R, bool GRAD>
__device__ __forceinline__ void load_pins(const NetsDev& n, const CellsDev& c, int s, int deg, int stride, int2 (&pos)[R],
                                          int (&cq)[R]) {
    int cl[R];
    int2 of[R];
#pragma unroll
    for (int i = 0; i < R; ++i) {
        cl[i] = -1;
        of[i] = make_int2(0, 0);
        cq[i] = -1;
        if (i * stride < deg) {
            cl[i] = n.pcell[s + i * stride];
            of[i] = n.poff[s + i * stride];
            if (GRAD) cq[i] = n.cpi[s + i * stride];
        }
    }
    // unguarded (an absent pin reads cell 0): a guard per read makes the compiler wait after each
#pragma unroll
    for (int i = 0; i < R; ++i) pos[i] = c.ctr[max(cl[i], 0)];
#pragma unroll
    for (int i = 0; i < R; ++i)
        if (i * stride < deg) pos[i] = pin_at(n, s + i * stride, cl[i], pos[i], of[i]);
}

// The four shifted exponentials of one pin (nesterov.rs:2354-2385) added into the net sums.
__device__ __forceinline__ void wa_pin(int2 p, int lx, int ly, int ux, int uy, float cx, float cy, float4& a, float4& b) {
    const float e0 = (float)(lx - p.x) * cx, e1 = (float)(p.x - ux) * cx;
    const float e2 = (float)(ly - p.y) * cy, e3 = (float)(p.y - uy) * cy;
    if (e0 > -300.0f) {
        const float v = fast_exp(e0);
        a.x += v;
        a.y += (float)p.x * v;
    }
    if (e1 > -300.0f) {
        const float v = fast_exp(e1);
        a.z += v;
        a.w += (float)p.x * v;
    }
    if (e2 > -300.0f) {
        const float v = fast_exp(e2);
        b.x += v;
        b.y += (float)p.y * v;
    }
    if (e3 > -300.0f) {
        const float v = fast_exp(e3);
        b.z += v;
        b.w += (float)p.y * v;
    }
}

// wa_pin that keeps the four exponentials in ev for the gradient pass (-1 where the term is
// absent, its exponent at or below -300; an exponential itself is never negative).
__device__ __forceinline__ void wa_pin_keep(int2 p, int lx, int ly, int ux, int uy, float cx, float cy, float4& a, float4& b, float4& ev) {
    const float e0 = (float)(lx - p.x) * cx, e1 = (float)(p.x - ux) * cx;
    const float e2 = (float)(ly - p.y) * cy, e3 = (float)(p.y - uy) * cy;
    ev = make_float4(-1.0f, -1.0f, -1.0f, -1.0f);
    if (e0 > -300.0f) {
        ev.x = fast_exp(e0);
        a.x += ev.x;
        a.y += (float)p.x * ev.x;
    }
    if (e1 > -300.0f) {
        ev.y = fast_exp(e1);
        a.z += ev.y;
        a.w += (float)p.x * ev.y;
    }
    if (e2 > -300.0f) {
        ev.z = fast_exp(e2);
        b.x += ev.z;
        b.y += (float)p.y * ev.z;
    }
    if (e3 > -300.0f) {
        ev.w = fast_exp(e3);
        b.z += ev.w;
        b.w += (float)p.y * ev.w;
    }
}

// getWireLengthGradientPinWA (nesterov.rs:2392): one exponential's term; the numerator and the
// quotient in double, the result a float.
__device__ __forceinline__ float wa_term_min(float S0, float S1, float v, float coef, float pc) {
    double num = (double)S0 * ((double)v * (1.0 - (double)(coef * pc))) + (double)(coef * v * S1);
    return (float)(num / (double)(S0 * S0));
}
__device__ __forceinline__ float wa_term_max(float S0, float S1, float v, float coef, float pc) {
    double num = (double)S0 * ((double)v * (1.0 + (double)(coef * pc))) - (double)(coef * v * S1);
    return (float)(num / (double)(S0 * S0));
}

// The quotient num / den of getWireLengthGradientPinWA's terms, correctly rounded in double as
// the division gives it, from the net's correctly rounded reciprocal y = 1 / den: q = num y, then
// one correction q + (num - q den) y (the residual exact by FMA), which is the correctly rounded
// quotient (Markstein's theorem; normal operands: den = S0^2 with 1 <= S0 <= the net's degree).
// One division per net and sum instead of one per pin and term.
struct WaRcp {
    double min_x, max_x, min_y, max_y;  // 1 / (double)(S0 * S0) of the four exponential sums
};
__device__ __forceinline__ WaRcp wa_rcp(float4 SX, float4 SY) {
    WaRcp r;
    r.min_x = 1.0 / (double)(SX.x * SX.x);
    r.max_x = 1.0 / (double)(SX.z * SX.z);
    r.min_y = 1.0 / (double)(SY.x * SY.x);
    r.max_y = 1.0 / (double)(SY.z * SY.z);
    return r;
}
__device__ __forceinline__ float quot(double num, float S0, double y) {
    const double den = (double)(S0 * S0);
    const double q = num * y;
    const double e = fma(-q, den, num);
    return (float)fma(e, y, q);
}
__device__ __forceinline__ float wa_term_min_r(float S0, float S1, float v, float coef, float pc, double y) {
    double num = (double)S0 * ((double)v * (1.0 - (double)(coef * pc))) + (double)(coef * v * S1);
    return quot(num, S0, y);
}
__device__ __forceinline__ float wa_term_max_r(float S0, float S1, float v, float coef, float pc, double y) {
    double num = (double)S0 * ((double)v * (1.0 + (double)(coef * pc))) - (double)(coef * v * S1);
    return quot(num, S0, y);
}

// One pin's weighted WA gradient (getWireLengthGradientPinWA, nesterov.rs:2392-2418, times the net
// weight) stored at its cell-pin index.
__device__ __forceinline__ void pin_grad(const NetsDev& n, int q, int2 p, int lx, int ly, int ux, int uy, float cx, float cy, float4 SX,
                                         float4 SY, float w, const WaRcp& rc) {
    if (q < 0) return;
    const float fpx = (float)p.x, fpy = (float)p.y;
    float gmin_x = 0.0f, gmax_x = 0.0f, gmin_y = 0.0f, gmax_y = 0.0f;
    const float e0 = (float)(lx - p.x) * cx, e1 = (float)(p.x - ux) * cx;
    const float e2 = (float)(ly - p.y) * cy, e3 = (float)(p.y - uy) * cy;
    if (e0 > -300.0f) gmin_x = wa_term_min_r(SX.x, SX.y, fast_exp(e0), cx, fpx, rc.min_x);
    if (e1 > -300.0f) gmax_x = wa_term_max_r(SX.z, SX.w, fast_exp(e1), cx, fpx, rc.max_x);
    if (e2 > -300.0f) gmin_y = wa_term_min_r(SY.x, SY.y, fast_exp(e2), cy, fpy, rc.min_y);
    if (e3 > -300.0f) gmax_y = wa_term_max_r(SY.z, SY.w, fast_exp(e3), cy, fpy, rc.max_y);
    float tx = gmin_x - gmax_x, ty = gmin_y - gmax_y;
    tx *= w;
    ty *= w;
    n.pgrad[q] = make_float2(tx, ty);
}

// pin_grad from the exponentials wa_pin_keep kept.
__device__ __forceinline__ void pin_grad_kept(const NetsDev& n, int q, int2 p, float cx, float cy, float4 SX, float4 SY, float w,
                                              float4 ev, const WaRcp& rc) {
    if (q < 0) return;
    const float fpx = (float)p.x, fpy = (float)p.y;
    float gmin_x = 0.0f, gmax_x = 0.0f, gmin_y = 0.0f, gmax_y = 0.0f;
    if (ev.x >= 0.0f) gmin_x = wa_term_min_r(SX.x, SX.y, ev.x, cx, fpx, rc.min_x);
    if (ev.y >= 0.0f) gmax_x = wa_term_max_r(SX.z, SX.w, ev.y, cx, fpx, rc.max_x);
    if (ev.z >= 0.0f) gmin_y = wa_term_min_r(SY.x, SY.y, ev.z, cy, fpy, rc.min_y);
    if (ev.w >= 0.0f) gmax_y = wa_term_max_r(SY.z, SY.w, ev.w, cy, fpy, rc.max_y);
    float tx = gmin_x - gmax_x, ty = gmin_y - gmax_y;
    tx *= w;
    ty *= w;
    n.pgrad[q] = make_float2(tx, ty);
}

// updateWireLengthForceWA (nesterov.rs:2325-2388), getHpwl (2297) and the pins' gradient terms,
// in two kernels.  SUMS false: the HPWL only.
//
// k_wa_big: one 1024-thread block per high-fanout net (num_small and up): the box and the sums
// reduced across the block, the sums in FP64 and rounded once to float.  Each thread holds up to
// kBigRegs of its pins in registers, loaded in one unrolled batch so that their gathers overlap.
constexpr int kBigBlock = 1024;
constexpr int kBigRegs = 8;
template <bool SUMS>
__device__ __forceinline__ void wa_big_body(NetsDev n, CellsDev c, float cx, float cy, unsigned long long* hpwl,
                                                      int write_rec, const unsigned bid_, const unsigned nblk_) {
    if (n.cg) {
        const int halt = n.cg->halt;
        cx = n.cg->wl_coef_x;
        cy = n.cg->wl_coef_y;
        if (halt) return;
    }
    __shared__ int shi[4][kBigBlock / 64];
    __shared__ double shd[8][kBigBlock / 64];
    __shared__ float shr[8];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int e = n.num_small + bid_;
    const int s = n.start[e], t = n.start[e + 1];
    const int rest = s + kBigRegs * kBigBlock;  // pins from here on are loaded again in each pass
    int2 pos[kBigRegs];
    int cq[kBigRegs];
    load_pins<kBigRegs, SUMS>(n, c, s + (int)threadIdx.x, t - s - (int)threadIdx.x, kBigBlock, pos, cq);
    int lx = INT_MAX, ly = INT_MAX, ux = INT_MIN, uy = INT_MIN;
#pragma unroll
    for (int i = 0; i < kBigRegs; ++i) {
        if (s + (int)threadIdx.x + i * kBigBlock < t) {
            lx = min(lx, pos[i].x);
            ly = min(ly, pos[i].y);
            ux = max(ux, pos[i].x);
            uy = max(uy, pos[i].y);
        }
    }
    for (int k = rest + threadIdx.x; k < t; k += blockDim.x) {
        const int2 p = pin_pos(n, c, k);
        lx = min(lx, p.x);
        ly = min(ly, p.y);
        ux = max(ux, p.x);
        uy = max(uy, p.y);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        lx = min(lx, __shfl_down(lx, o, 64));
        ly = min(ly, __shfl_down(ly, o, 64));
        ux = max(ux, __shfl_down(ux, o, 64));
        uy = max(uy, __shfl_down(uy, o, 64));
    }
    if (lane == 0) {
        shi[0][w] = lx;
        shi[1][w] = ly;
        shi[2][w] = ux;
        shi[3][w] = uy;
    }
    __syncthreads();
    lx = shi[0][0];
    ly = shi[1][0];
    ux = shi[2][0];
    uy = shi[3][0];
    for (int i = 1; i < kBigBlock / 64; ++i) {
        lx = min(lx, shi[0][i]);
        ly = min(ly, shi[1][i]);
        ux = max(ux, shi[2][i]);
        uy = max(uy, shi[3][i]);
    }
    if (SUMS) {
        double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        auto add = [&](int2 p) {
            float4 a = make_float4(0.f, 0.f, 0.f, 0.f), b = make_float4(0.f, 0.f, 0.f, 0.f);
            wa_pin(p, lx, ly, ux, uy, cx, cy, a, b);
            acc[0] += a.x;
            acc[1] += a.y;
            acc[2] += a.z;
            acc[3] += a.w;
            acc[4] += b.x;
            acc[5] += b.y;
            acc[6] += b.z;
            acc[7] += b.w;
        };
#pragma unroll
        for (int i = 0; i < kBigRegs; ++i)
            if (s + (int)threadIdx.x + i * kBigBlock < t) add(pos[i]);
        for (int k = rest + threadIdx.x; k < t; k += blockDim.x) add(pin_pos(n, c, k));
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            double v = wave_sum(acc[q]);
            if (lane == 0) shd[q][w] = v;
        }
        __syncthreads();
        if (threadIdx.x < 8) {
            double v = 0.0;
            for (int i = 0; i < kBigBlock / 64; ++i) v += shd[threadIdx.x][i];
            shr[threadIdx.x] = (float)v;
        }
        __syncthreads();
        const float4 SX = make_float4(shr[0], shr[1], shr[2], shr[3]);
        const float4 SY = make_float4(shr[4], shr[5], shr[6], shr[7]);
        if (threadIdx.x == 0 && write_rec) {
            NetRec& R = n.rec[e];
            R.box = make_int4(lx, ly, ux, uy);
            R.sx = SX;
            R.sy = SY;
        }
        const float wt = n.weight[e];
        const WaRcp rc = wa_rcp(SX, SY);
#pragma unroll
        for (int i = 0; i < kBigRegs; ++i) {
            const int k = s + threadIdx.x + i * kBigBlock;
            if (k < t) pin_grad(n, cq[i], pos[i], lx, ly, ux, uy, cx, cy, SX, SY, wt, rc);
        }
        for (int k = rest + threadIdx.x; k < t; k += blockDim.x)
            pin_grad(n, n.cpi[k], pin_pos(n, c, k), lx, ly, ux, uy, cx, cy, SX, SY, wt, rc);
    }
    if (threadIdx.x == 0 && ux >= lx) atomicAdd(hpwl, (unsigned long long)((long long)ux - lx + (long long)uy - ly));
}

template <bool SUMS>
__global__ void __launch_bounds__(kBigBlock) k_wa_big(NetsDev n, CellsDev c, float cx, float cy, unsigned long long* hpwl,
                                                      int write_rec) {
    wa_big_body<SUMS>(n, c, cx, cy, hpwl, write_rec, blockIdx.x, gridDim.x);
}

// k_wa_tiles: the small nets (below num_small) in tiles of consecutive nets (the nets whose first
// pin lies in one range of kTileSpan pins: at most kTilePins pins; XCD x takes a contiguous
// eighth of the tiles).  A block first stages its
// tile's pin positions (and cell-pin indices) in LDS, pin-parallel: the pin records are read
// coalesced and every cell-centre gather of the tile is in flight at once.  Then one net per
// thread reads its pins from LDS in the net's order: the box, the four shifted exponentials
// summed in float in pin order (the reference's sums, bit for bit), the first kKeepPins pins'
// exponentials kept in registers for the gradient pass, each pin's weighted gradient term stored
// at its cell-pin index.  The nets of a tile have close degrees (nets are sorted by degree within
// their window), so a wave's lanes loop alike.
constexpr int kKeepPins = 8;  // pins whose four exponentials are kept in registers
constexpr int kTileStage = kTilePins / kBlock;
template <bool SUMS>
__device__ __forceinline__ void wa_tiles_body(NetsDev n, CellsDev c, float cx, float cy, unsigned long long* hpwl,
                                                     int write_rec, const unsigned bid_, const unsigned nblk_) {
    if (n.cg) {
        const int halt = n.cg->halt;
        cx = n.cg->wl_coef_x;
        cy = n.cg->wl_coef_y;
        if (halt) return;
    }
    __shared__ int2 spos[kTilePins];
    __shared__ int scq[SUMS ? kTilePins : 1];
    __shared__ unsigned long long shh[kBlock / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int x = bid_ % kXcds;
    const int t = x * n.tile_blocks + (int)(bid_ / kXcds);
    unsigned long long h = 0;
    if (t < n.num_tiles) {
        const int e0 = n.tile[t], e1 = n.tile[t + 1];
        const int s0 = n.start[e0], np = n.start[e1] - s0;
        {
            int cl[kTileStage], cq[kTileStage];
            int2 of[kTileStage], m[kTileStage];
#pragma unroll
            for (int i = 0; i < kTileStage; ++i) {
                const int j = (int)threadIdx.x + i * kBlock;
                cl[i] = -1;
                cq[i] = -1;
                of[i] = make_int2(0, 0);
                if (j < np) {
                    cl[i] = n.pcell[s0 + j];
                    of[i] = n.poff[s0 + j];
                    if (SUMS) cq[i] = n.cpi[s0 + j];
                }
            }
            // unguarded (an absent pin reads cell 0): a guard per read makes the compiler wait after each
#pragma unroll
            for (int i = 0; i < kTileStage; ++i) m[i] = c.ctr[max(cl[i], 0)];
#pragma unroll
            for (int i = 0; i < kTileStage; ++i) {
                const int j = (int)threadIdx.x + i * kBlock;
                if (j < np) {
                    spos[j] = pin_at(n, s0 + j, cl[i], m[i], of[i]);
                    if (SUMS) scq[j] = cq[i];
                }
            }
        }
        __syncthreads();
        for (int e = e0 + (int)threadIdx.x; e < e1; e += kBlock) {
            const int s = n.start[e] - s0, deg = n.start[e + 1] - n.start[e];
            int lx = INT_MAX, ly = INT_MAX, ux = INT_MIN, uy = INT_MIN;
            for (int k = 0; k < deg; ++k) {
                const int2 p = spos[s + k];
                lx = min(lx, p.x);
                ly = min(ly, p.y);
                ux = max(ux, p.x);
                uy = max(uy, p.y);
            }
            if (SUMS) {
                float4 a = make_float4(0.f, 0.f, 0.f, 0.f), b = make_float4(0.f, 0.f, 0.f, 0.f);
                float4 ev[kKeepPins];
#pragma unroll
                for (int k = 0; k < kKeepPins; ++k)
                    if (k < deg) wa_pin_keep(spos[s + k], lx, ly, ux, uy, cx, cy, a, b, ev[k]);
                for (int k = kKeepPins; k < deg; ++k) wa_pin(spos[s + k], lx, ly, ux, uy, cx, cy, a, b);
                if (write_rec) {
                    NetRec& R = n.rec[e];
                    R.box = make_int4(lx, ly, ux, uy);
                    R.sx = a;
                    R.sy = b;
                }
                const float wt = n.weight[e];
                const WaRcp rc = wa_rcp(a, b);
#pragma unroll
                for (int k = 0; k < kKeepPins; ++k)
                    if (k < deg) pin_grad_kept(n, scq[s + k], spos[s + k], cx, cy, a, b, wt, ev[k], rc);
                for (int k = kKeepPins; k < deg; ++k) pin_grad(n, scq[s + k], spos[s + k], lx, ly, ux, uy, cx, cy, a, b, wt, rc);
            }
            if (ux >= lx) h += (unsigned long long)((long long)ux - lx + (long long)uy - ly);
        }
    }
    h = wave_sum(h);
    if (lane == 0) shh[w] = h;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tot = 0;
        for (int i = 0; i < kBlock / 64; ++i) tot += shh[i];
        if (tot) atomicAdd(hpwl, tot);
    }
}

template <bool SUMS>
__global__ void __launch_bounds__(kBlock) k_wa_tiles(NetsDev n, CellsDev c, float cx, float cy, unsigned long long* hpwl,
                                                     int write_rec) {
    wa_tiles_body<SUMS>(n, c, cx, cy, hpwl, write_rec, blockIdx.x, gridDim.x);
}

constexpr int kGradPins = 6;
// updateGradients (nesterov.rs:2455-2489) per cell: the WA gradient over the cell's pins in
// their order (each pin's weighted term from the WA kernel); the density gradient over the bins its box
// covers, x outer, y inner (2434); the preconditioned sum; FP64 partial sums of |wl|, |den|,
// and, for a trial, the squared coordinate and gradient distances (nesterovUpdateStepLength,
// 2755-2756) to a_slp/a_sum.
// (six waves per SIMD: 76 registers, no spill)
__device__ __forceinline__ void gradient_body(RegionDev r, CellsDev c, NetsDev nt, float cx, float cy, float penalty, float2* wl_out, float2* den_out,
                           float2* sum_out, const float2* a_slp, const float2* b_slp, const float2* a_sum, double* partials,
                           int nan_inject, const unsigned bid_, const unsigned nblk_) {
    int n = r.cell_end - r.cell_begin;
    if (r.ctl) {
        // a device-driven trial: the penalty and the rotated slots from the control record's line
        // (a skipped region: no cells, its partials zero and unread)
        const CtlHot h = r.ctl->hot;
        penalty = h.density_penalty;
        sum_out = h.next_sum;
        a_slp = h.cur_slp;
        b_slp = h.next_slp;
        a_sum = h.cur_sum;
        wl_out = nullptr;
        den_out = nullptr;
        nan_inject = (h.flags >> 1) & 1;
        if (h.flags & 1) n = 0;
    }
    __shared__ double sh[kBlock / 64];
    __shared__ float2 spg[kGradStage];
    double s_wl = 0.0, s_den = 0.0, s_c = 0.0, s_g = 0.0;
    // the block's cells in groups of kBlock consecutive cells (the trip count is the block's own)
    for (int base = bid_ * blockDim.x; base < n; base += nblk_ * blockDim.x) {
        // the group's pin terms are one contiguous range (cells in order, each cell's pins in
        // order): read coalesced into LDS when it fits
        const int qb = c.pin_start[r.cell_begin + base];
        const int qe = c.pin_start[r.cell_begin + min(base + kBlock, n)];
        const bool staged = qe - qb <= min(r.grad_stage, kGradStage);
        // The cell's own reads and its 3 x 3 bin window of the field are issued first, so that
        // they are in flight with the staging of the group's pin terms (a lane past the region's
        // end reads its last cell and stores nothing); indices are clamped into the arrays, with
        // no guards, so that the compiler does not wait after each read; cells with more pins or
        // a wider span finish in the loops below.
        const int k = base + (int)threadIdx.x;
        const bool live = k < n;
        const int g = r.cell_begin + min(k, n - 1);
        const int4 d = c.dbox[g];
        const int q0 = c.pin_start[g], q1 = c.pin_start[g + 1];
        const float scale = c.dscale[g];
        const float2 pr = c.pre[g];
        int x0, x1, y0, y1;
        bin_span(d.x, d.z, r.lx, r.size_x, r.cnt_x, x0, x1);
        bin_span(d.y, d.w, r.ly, r.size_y, r.cnt_y, y0, y1);
        int ex[4], ey[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            ex[i] = bin_edge(r.bin_x, r.edges_rounded, r.lx, r.size_x, r.cnt_x, x0 + i);
            ey[i] = bin_edge(r.bin_y, r.edges_rounded, r.ly, r.size_y, r.cnt_y, y0 + i);
        }
        float2 wf[9];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const size_t kb = (size_t)min(y0 + j, r.cnt_y - 1) * r.cnt_x + min(x0 + i, r.cnt_x - 1);
                wf[3 * i + j] = r.field[kb];
            }
        }
        __syncthreads();
        if (staged) {
            // four reads in flight per thread before their stores; indices past the range clamp
            // to its last term (read and stored again, the same value), so there are no guards
            const int cnt = qe - qb;
            for (int i0 = (int)threadIdx.x; i0 < cnt; i0 += 4 * kBlock) {
                float2 v[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) v[u] = nt.pgrad[qb + min(i0 + u * kBlock, cnt - 1)];
#pragma unroll
                for (int u = 0; u < 4; ++u) spg[min(i0 + u * kBlock, cnt - 1)] = v[u];
            }
        }
        __syncthreads();
        if (!live) continue;
        float2 t[kGradPins];
        // a pinless cell (q0 == q1) clamps to a term inside the staged range (or the array), so
        // every read stays in bounds; its values are not used
        const int qlast = staged ? max(q1 - 1, qb) : max(q1 - 1, 0);
        if (staged) {
#pragma unroll
            for (int j = 0; j < kGradPins; ++j) t[j] = spg[min(q0 + j, qlast) - qb];
        } else {
#pragma unroll
            for (int j = 0; j < kGradPins; ++j) t[j] = nt.pgrad[min(q0 + j, qlast)];
        }
        // the WA gradient: the pins' terms in the cell's pin order
        float wx = 0.0f, wy = 0.0f;
#pragma unroll
        for (int j = 0; j < kGradPins; ++j) {
            if (q0 + j < q1) {
                wx += t[j].x;
                wy += t[j].y;
            }
        }
        for (int q = q0 + kGradPins; q < q1; ++q) {
            const float2 u = staged ? spg[q - qb] : nt.pgrad[q];
            wx += u.x;
            wy += u.y;
        }
        // density gradient, x outer, y inner
        float fx = 0.0f, fy = 0.0f;
        if (x1 - x0 <= 3 && y1 - y0 <= 3) {
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                if (i >= x1 - x0) break;
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    if (j >= y1 - y0) break;
                    float a = overlap_area(ex[i], ey[j], ex[i + 1], ey[j + 1], d) * scale;
                    fx += a * wf[3 * i + j].x;
                    fy += a * wf[3 * i + j].y;
                }
            }
        } else {
            for (int x = x0; x < x1; ++x) {
                int bx0 = r.bin_x[x], bx1 = r.bin_x[x + 1];
                for (int y = y0; y < y1; ++y) {
                    float a = overlap_area(bx0, r.bin_y[y], bx1, r.bin_y[y + 1], d) * scale;
                    size_t kb = (size_t)y * r.cnt_x + x;
                    const float2 f = r.field[kb];
                    fx += a * f.x;
                    fy += a * f.y;
                }
            }
        }
        float sx = wx + penalty * fx;
        float sy = wy + penalty * fy;
        float pre = rmax(pr.x + penalty * pr.y, 1.0f);
        sx /= pre;
        sy /= pre;
        if (wl_out) {
            wl_out[k] = make_float2(wx, wy);
            den_out[k] = make_float2(fx, fy);
        }
        sum_out[k] = make_float2(sx, sy);
        s_wl += (double)fabsf(wx) + (double)fabsf(wy);
        s_den += (double)fabsf(fx) + (double)fabsf(fy);
        if (a_slp) {
            float2 pa = a_slp[k], pb = b_slp[k], ga = a_sum[k];
            float dx = pa.x - pb.x, dy = pa.y - pb.y;
            float gx = ga.x - sx, gy = ga.y - sy;
            s_c += (double)(dx * dx) + (double)(dy * dy);
            s_g += (double)(gx * gx) + (double)(gy * gy);
        }
    }
    double t0 = block_sum(s_wl, sh);
    double t1 = block_sum(s_den, sh);
    double t2 = block_sum(s_c, sh);
    double t3 = block_sum(s_g, sh);
    if (threadIdx.x == 0) {
        if (nan_inject) t0 = __longlong_as_double(0x7ff8000000000000ll);
        partials[2 * kPartials + bid_] = t0;
        partials[3 * kPartials + bid_] = t1;
        partials[4 * kPartials + bid_] = t2;
        partials[5 * kPartials + bid_] = t3;
    }
}

__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6, 8))) k_gradient(RegionDev r, CellsDev c, NetsDev nt, float cx, float cy, float penalty, float2* wl_out, float2* den_out,
                           float2* sum_out, const float2* a_slp, const float2* b_slp, const float2* a_sum, double* partials,
                           int nan_inject) {
    gradient_body(r, c, nt, cx, cy, penalty, wl_out, den_out, sum_out, a_slp, b_slp, a_sum, partials, nan_inject, blockIdx.x, gridDim.x);
}

// Squared distances between coordinate sets (getDistance's sums, nesterov.rs:319) as partials 4
// (a - b) and 5 (c - d).
__global__ void k_distance(int n, const float2* a, const float2* b, const float2* c, const float2* d, double* partials) {
    __shared__ double sh[kBlock / 64];
    double s0 = 0.0, s1 = 0.0;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        float2 p = a[k], q = b[k], u = c[k], v = d[k];
        float dx = p.x - q.x, dy = p.y - q.y, gx = u.x - v.x, gy = u.y - v.y;
        s0 += (double)(dx * dx) + (double)(dy * dy);
        s1 += (double)(gx * gx) + (double)(gy * gy);
    }
    double t0 = block_sum(s0, sh), t1 = block_sum(s1, sh);
    if (threadIdx.x == 0) {
        partials[4 * kPartials + blockIdx.x] = t0;
        partials[5 * kPartials + blockIdx.x] = t1;
    }
}

// updateInitialPrevSLPCoordi (nesterov.rs:2583).
__global__ void k_initial_prev(RegionDev r, CellsDev c, const float2* cur_slp, const float2* cur_sum, float coef, float2* prev_slp) {
    int n = r.cell_end - r.cell_begin;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        int g = r.cell_begin + k;
        float2 a = cur_slp[k];
        if (c.flags[g] & 4u) {
            prev_slp[k] = a;
            continue;
        }
        int4 d = c.dbox[g];
        float px = a.x - coef * cur_sum[k].x, py = a.y - coef * cur_sum[k].y;
        float hx = (float)(d.z - d.x) / 2.0f, hy = (float)(d.w - d.y) / 2.0f;
        prev_slp[k] = make_float2(inside(px, hx, (float)r.lx, (float)r.ux), inside(py, hy, (float)r.ly, (float)r.uy));
    }
}

// The finish: each quantity's block partials summed in index order by one block (a fixed
// per-thread stride then a fixed tree).
// One 1024-thread block: the six sums at once (each in a fixed order: thread t adds partials t,
// t + 1024, ..., then the waves' totals in wave order).
// ---------------------------------------------------------------------------------------------
// Device-driven trials: the loop's decisions after a trial, on the device (one thread).  Every
// expression is the host's (nesterov.cpp: gradients_read, update_step_length, do_back_tracking,
// adjust_phi, nb_/np_update_next_iter, update_wirelength_coef, phi_coef, the momentum of loop_step,
// maybe_resort's bookkeeping), in the same float and double arithmetic, so that the two modes give
// the same rows; the events the host handles halt the device before any of their effects.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ bool ctl_bad(float v) { return isnan(v) || isinf(v); }

__device__ void ctl_halt(CtlHead& g, int event, int stage) {
    g.halt = event;  // (ctl_refresh then sets every region's skip)
    g.stage = stage;
}

__device__ void ctl_wl_coef(CtlHead& g, float overflow) {
#pragma clang fp contract(off)
    float c;
    if (overflow > 1.0f)
        c = (float)0.1;
    else if ((double)overflow < 0.1)
        c = 10.0f;
    else
        c = (float)(1.0 / pow(10.0, ((double)overflow - 0.1) * 20.0 / 9.0 - 1.0));
    g.wl_coef_x = c * g.base_wl_coef;
    g.wl_coef_y = c * g.base_wl_coef;
}

// g and the sums (every region's, then the tail) staged in LDS; history and progress rows appended
// to G's rings; nstep / ncd: LDS scratch per region
__device__ void ctl_trial(CtlHead& g, const Sums* sums, CtlGlobal* G, float* nstep, float* ncd) {
#pragma clang fp contract(off)
    const Sums* tail = sums + g.nreg;
    const int R = g.nreg;
    // gradients_read: a zero WA sum or a non-finite sum is the host's (nothing changed yet); a
    // converged region's are its last values, as on the host
    for (int b = 0; b < R; ++b) {
        const CtlRegion& q = g.reg[b];
        const float wl = q.is_converged ? q.wl_grad_sum : (float)sums[b].wl_abs;
        const float den = q.is_converged ? q.density_grad_sum : (float)sums[b].den_abs;
        if (wl == 0.0f && g.recursion_wl < 10) return ctl_halt(g, CTL_ZERO_WL, 1);
        if (g.iter == g.zero_wl_at && !q.is_converged) return ctl_halt(g, CTL_ZERO_WL, 1);
        if (ctl_bad(wl) || ctl_bad(den)) return ctl_halt(g, CTL_NAN, 1);
    }
    // nesterovUpdateStepLength of every live region, checked before any is kept
    int ok = 0;
    for (int b = 0; b < R; ++b) {
        const CtlRegion& q = g.reg[b];
        if (q.is_converged) {
            ++ok;
            continue;
        }
        const double n2 = 2.0 * (double)q.n;
        const float cd = q.n ? (float)sqrt(sums[b].coord_sq / n2) : 0.0f;
        const float gd = q.n ? (float)sqrt(sums[b].grad_sq / n2) : 0.0f;
        const float nw = cd / gd;
        if (ctl_bad(nw) || (g.step_nan_from >= 0 && g.iters_run >= g.step_nan_from)) return ctl_halt(g, CTL_STEP, 1);
        ncd[b] = cd;
        if ((double)nw > (double)q.step_length * 0.95) {
            nstep[b] = nw;
        } else if ((double)nw < 0.01) {
            nstep[b] = (float)0.01;
        } else {
            nstep[b] = nw;
            ++ok;
        }
    }
    for (int b = 0; b < R; ++b) {
        CtlRegion& q = g.reg[b];
        if (q.is_converged) continue;
        q.wl_grad_sum = (float)sums[b].wl_abs;
        q.density_grad_sum = (float)sums[b].den_abs;
        q.overflow = (long long)sums[b].overflow;
        q.overflow_unscaled = (long long)sums[b].overflow_unscaled;
        q.step_length = nstep[b];
        q.coordi_distance = ncd[b];
    }
    g.cur_hpwl = (long long)tail->hpwl;
    g.trials += 1;
    if (ok == R && g.trials < 10) return;  // every step shrank: another trial
    // the iteration's end
    const int iter = g.iter;
    const int trials = g.trials;
    g.total_trials += trials;
    for (int b = 0; b < R; ++b) {  // nesterovAdjustPhi
        CtlRegion& q = g.reg[b];
        if (q.is_converged) continue;
        if (!q.max_phi_coef_changed && q.sum_overflow_unscaled < 0.35f) {
            q.max_phi_coef_changed = 1;
            q.max_phi_coef = (float)((double)q.max_phi_coef * 0.99);
        }
        if (q.max_phi_coef <= 1.0f) q.max_phi_coef = 1.01f;
    }
    float total = 0.0f, total_unscaled = 0.0f;
    for (int b = 0; b < R; ++b) {  // NesterovBase::updateNextIter
        CtlRegion& q = g.reg[b];
        if (!q.is_converged) {
            q.rot = q.rot == 2 ? 0 : q.rot + 1;
            q.cnp ^= 1;
            const float fraction = (float)iter / (float)g.max_iter;
            const float a = (float)q.nia, c = fraction * (float)q.non_place_insts_area * 0.05f;
            const float denom = a > c ? a : c;
            q.sum_overflow = (float)q.overflow / denom;
            q.sum_overflow_unscaled = (float)q.overflow_unscaled / denom;
            const long long hpwl = g.cur_hpwl;
            if (iter == 0 || iter % 10 == 0) {
                float pct = 0.0f;
                if (q.prev_reported_hpwl != 0)
                    pct = (float)((double)(hpwl - q.prev_reported_hpwl) / (double)q.prev_reported_hpwl * 100.0);
                if (g.nlog < kCtlRing) {
                    CtlLog& L = G->log[g.nlog++];
                    L.iter = iter;
                    L.region = b;
                    L.overflow = q.sum_overflow_unscaled;
                    L.pct = pct;
                    L.penalty = q.density_penalty;
                    L.hpwl = hpwl;
                }
                q.prev_reported_hpwl = hpwl;
                q.prev_reported_overflow_unscaled = q.sum_overflow_unscaled;
            }
            const float scaled = (float)(hpwl - q.prev_hpwl) / g.reference_hpwl;
            const float m = q.max_phi_coef;
            const float v = scaled < 0.0f ? m : (float)((double)m * pow((double)m, -(double)scaled));
            const float phi = g.min_phi > v ? g.min_phi : v;
            q.density_penalty *= phi;
            q.prev_hpwl = hpwl;
            q.peak_coordi_distance = q.peak_coordi_distance > q.coordi_distance ? q.peak_coordi_distance : q.coordi_distance;
            if (iter > 50 && q.min_sum_overflow > q.sum_overflow_unscaled) {
                q.min_sum_overflow = q.sum_overflow_unscaled;
                q.hpwl_with_min_sum_overflow = (float)q.prev_hpwl;
            }
        }
        total += q.sum_overflow;
        total_unscaled += q.sum_overflow_unscaled;
    }
    const float nr = (float)R;
    g.average_overflow = total / nr;
    g.average_overflow_unscaled = total_unscaled / nr;
    ctl_wl_coef(g, g.average_overflow);
    if (!g.disable_revert) {
        const long long hpwl = g.cur_hpwl;
        if (hpwl < g.min_hpwl && (double)g.average_overflow_unscaled <= 0.25) {
            g.min_hpwl = hpwl;
            g.diverge_snapshot_overflow = g.average_overflow_unscaled;
            g.diverge_snapshot_iter = iter + 1;
            g.is_min_hpwl = 1;
        } else {
            g.is_min_hpwl = 0;
        }
    }
    if (g.nhist < kCtlRing) {  // push_history
        CtlHist& H = G->hist[g.nhist++];
        H.iter = iter;
        H.trials = trials;
        H.overflow = g.average_overflow_unscaled;
        H.density_penalty = g.reg[0].density_penalty;
        H.step_length = g.reg[0].step_length;
        H.wl_coef = g.wl_coef_x;
        H.hpwl = g.reg[0].prev_hpwl;
    }
    g.iters_run += 1;
    g.trials = 0;
    // the host takes over for isDiverged's effects, a new snapshot and a convergence
    for (int b = 0; b < R; ++b) {  // checkDivergence's rules (every region, as on the host)
        const CtlRegion& q = g.reg[b];
        bool d = q.sum_overflow_unscaled < 0.2f && q.sum_overflow_unscaled - q.min_sum_overflow >= 0.02f &&
                 q.hpwl_with_min_sum_overflow * 1.2f < (float)q.prev_hpwl;
        if (q.prev_reported_overflow_unscaled > 0.0f && q.prev_reported_hpwl > 0) {
            const float change = q.sum_overflow_unscaled - q.prev_reported_overflow_unscaled;
            const float increase = (float)(q.prev_hpwl - q.prev_reported_hpwl) / (float)q.prev_reported_hpwl;
            d = d || (change >= 0.05f && increase >= 0.25f);
        }
        if (d || iter == g.diverge_at || iter == g.diverge_at2) return ctl_halt(g, CTL_DIVERGE, 2);
    }
    if (g.is_min_hpwl && !g.disable_revert) return ctl_halt(g, CTL_SNAPSHOT, 2);
    for (int b = 0; b < R; ++b)
        if (!g.reg[b].is_converged && !(g.reg[b].sum_overflow_unscaled > g.overflow_target)) return ctl_halt(g, CTL_CONVERGE, 2);
    // maybe_resort's bookkeeping; a due re-sort is the host's
    if (g.resort_enabled) {
        if (g.resort_from >= 0 && iter >= g.resort_from) return ctl_halt(g, CTL_RESORT, 2);
        const unsigned long long ga = (unsigned long long)(unsigned)tail->pad[0] | ((unsigned long long)(unsigned)tail->pad[1] << 32);
        const double rate = double(ga - g.last_global_adds) / double(trials > 1 ? trials : 1);
        double base = g.adds_base;
        if (iter == g.last_resort_iter + 1 || base == 0.0) base = rate;
        if (rate > (double)g.resort_growth * base && iter - g.last_resort_iter >= g.resort_gap) return ctl_halt(g, CTL_RESORT, 2);
        g.last_global_adds = ga;
        g.adds_base = base;
    }
    g.iter = iter + 1;
    if (g.iter >= g.max_iter || g.iter >= g.iter_stop || g.nhist >= kCtlRing || g.nlog + R > kCtlRing)
        return ctl_halt(g, CTL_STOP, 3);
    // the next iteration's momentum
    const float prev_a = g.cur_a;
    g.cur_a = (float)((1.0 + sqrt(4.0 * (double)prev_a * (double)prev_a + 1.0)) * 0.5);
    g.coeff = (float)(((double)prev_a - 1.0) / (double)g.cur_a);
}

// The control of a device-driven trial, run by every thread of k_finish's block once the record's
// header and regions (st) and every region's sums (ss) are in LDS: the decisions by one thread, the
// record written back, the host signalled.
__device__ void ctl_decide(const CtlArgs& ca, CtlHead& st, Sums* ss, float* nstep, float* ncd, int words, unsigned long long seq) {
    __shared__ int run;
    if (threadIdx.x == 0) {
        run = st.halt == 0;
        if (run) {
            ctl_trial(st, ss, ca.g, nstep, ncd);
            ctl_refresh(st);
        }
    }
    __syncthreads();
    unsigned long long* src = reinterpret_cast<unsigned long long*>(&ca.g->h);
    const unsigned long long* dst = reinterpret_cast<const unsigned long long*>(&st);
    if (run && (int)threadIdx.x < words) src[threadIdx.x] = dst[threadIdx.x];
    if (threadIdx.x == 0) {
        // the next trial's WA kernels add into a cleared counter; a halted trial's HPWL stays for the host
        if (run && !st.halt) ca.tail->hpwl = 0ull;
        const unsigned long long q = seq + 1ull;
        *ca.dev_seq = q;
        ca.host_seq[1] = (unsigned long long)(unsigned)st.halt;
        __hip_atomic_store(ca.host_seq, q, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

constexpr int kFinishBlock = 1024;
__device__ __forceinline__ void finish_body(const double* partials, int nb_bins, int nb_cells, Sums* out, const unsigned long long* hpwl,
                                            Sums* host_out, const Sums* dev_tail, Sums* host_tail, unsigned long long* dev_seq,
                                            unsigned long long* host_seq, Sums* lds_out) {
    __shared__ double sh[6][kFinishBlock / 64];
    // thread 0's scalars read first, so that their latency overlaps the partials'
    unsigned long long seq = 0, h = 0, tail_h = 0, tail_p0 = 0, tail_p1 = 0;
    if (threadIdx.x == 0) {
        if (host_seq) seq = *dev_seq;
        if (hpwl) h = *hpwl;
        if (host_tail) {
            tail_h = dev_tail->hpwl;
            tail_p0 = dev_tail->pad[0];
            tail_p1 = dev_tail->pad[1];
        }
    }
    // each thread's two partials of every quantity (slots tid and tid + kFinishBlock) read at
    // once, then added in that order: the sums of a strided loop, without a wait after each
    // read.  Only the written slots [0, cnt) of a quantity are read (indices clamped to cnt - 1,
    // none for a quantity with no partials): a caller's buffer may hold only the quantities it
    // counts (the bins' two, with no cells).
    static_assert(kPartials <= 2 * kFinishBlock, "two partial slots per thread and quantity");
    double a0[6], a1[6];
    const int i0 = (int)threadIdx.x, i1 = (int)threadIdx.x + kFinishBlock;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const int cnt = q < 2 ? nb_bins : nb_cells;
        a0[q] = 0.0;
        a1[q] = 0.0;
        if (cnt > 0) {
            a0[q] = partials[q * kPartials + min(i0, cnt - 1)];
            a1[q] = partials[q * kPartials + min(i1, cnt - 1)];
        }
    }
    double v[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const int cnt = q < 2 ? nb_bins : nb_cells;
        double s = 0.0;
        if (i0 < cnt) s += a0[q];
        if (i1 < cnt) s += a1[q];
        v[q] = wave_sum(s);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < 6; ++q) sh[q][w] = v[q];
    __syncthreads();
    if (threadIdx.x < 6) {
        double t = 0.0;
        for (int i = 0; i < kFinishBlock / 64; ++i) t += sh[threadIdx.x][i];
        double* o = threadIdx.x == 0 ? &out->overflow
                    : threadIdx.x == 1 ? &out->overflow_unscaled
                    : threadIdx.x == 2 ? &out->wl_abs
                    : threadIdx.x == 3 ? &out->den_abs
                    : threadIdx.x == 4 ? &out->coord_sq
                                       : &out->grad_sq;
        if ((threadIdx.x < 2 && nb_bins > 0) || (threadIdx.x >= 2 && nb_cells > 0)) {
            *o = t;
            // the same value straight into the host's pinned copy (no copy node after the trial)
            if (host_out) *(reinterpret_cast<double*>(host_out) + (o - &out->overflow)) = t;
            if (lds_out) *(reinterpret_cast<double*>(lds_out) + (o - &out->overflow)) = t;
        }
    }
    if (threadIdx.x == 0 && hpwl) {
        out->hpwl = h;
        if (host_out) host_out->hpwl = h;
        if (lds_out) lds_out->hpwl = h;
    }
    if (threadIdx.x == 0 && host_tail) {
        // the HPWL counter and the scatter's global-add count
        host_tail->hpwl = tail_h;
        host_tail->pad[0] = tail_p0;
        host_tail->pad[1] = tail_p1;
    }
    if (host_seq) {
        // the trial's last finishing reduction: once every host write above is visible system-wide,
        // the next sequence number tells the waiting host that the statistics are complete
        if (threadIdx.x < 6) __threadfence_system();
        __syncthreads();
        if (threadIdx.x == 0) {
            // (the host reads only host_seq: the device copy needs no ordering against it).  The
            // signal is a system-scope release store, so every write this block made before the
            // barrier is visible to the host before the new number is.
            const unsigned long long q = seq + 1ull;
            *dev_seq = q;
            __hip_atomic_store(host_seq, q, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

__global__ void __launch_bounds__(kFinishBlock) k_finish(const double* partials, int nb_bins, int nb_cells, Sums* out,
                                                         const unsigned long long* hpwl, Sums* host_out,
                                                         const Sums* dev_tail, Sums* host_tail, unsigned long long* dev_seq,
                                                         unsigned long long* host_seq, const int* skip, CtlArgs ctl) {
    __shared__ CtlHead st;
    __shared__ Sums ss[kCtlMaxRegions + 1];
    __shared__ float nstep[kCtlMaxRegions], ncd[kCtlMaxRegions];
    const bool skipped = skip && (*skip & 1);
    if (skipped && !ctl.g) return;
    // the control's staging loads are issued first (one word per thread: the header and regions
    // take at most 8 x 84 words, the sums 8 x 17), so that they fly beside the reduction's
    int words = 0, swords = 0;
    unsigned long long v0 = 0, v1 = 0, seq = 0;
    if (ctl.g) {
        words = int((offsetof(CtlHead, reg) + size_t(ctl.nreg) * sizeof(CtlRegion)) / 8);
        swords = int(size_t(ctl.nreg + 1) * sizeof(Sums) / 8);
        if ((int)threadIdx.x < words) v0 = reinterpret_cast<const unsigned long long*>(&ctl.g->h)[threadIdx.x];
        if ((int)threadIdx.x < swords) v1 = reinterpret_cast<const unsigned long long*>(ctl.sums)[threadIdx.x];
        if (threadIdx.x == 0) seq = *ctl.dev_seq;
    }
    if (!skipped)
        finish_body(partials, nb_bins, nb_cells, out, hpwl, host_out, dev_tail, host_tail, dev_seq, host_seq,
                    ctl.g ? &ss[ctl.nreg - 1] : nullptr);
    if (ctl.g) {
        // the staged words into LDS, but not those of this region's sums that the reduction wrote
        const int w = (int)threadIdx.x - 8 * (ctl.nreg - 1);
        const bool fresh = !skipped && w >= 0 && w < 8 && ((w < 2 && nb_bins > 0) || (w >= 2 && w < 6 && nb_cells > 0) || (w == 6 && hpwl));
        if ((int)threadIdx.x < words) reinterpret_cast<unsigned long long*>(&st)[threadIdx.x] = v0;
        if ((int)threadIdx.x < swords && !fresh) reinterpret_cast<unsigned long long*>(ss)[threadIdx.x] = v1;
        __syncthreads();
        ctl_decide(ctl, st, ss, nstep, ncd, words, seq);
    }
}

// nesterov.rs:2850-2859: a locked instance cell's next slots take its current values.
__global__ void k_lock_fix(RegionDev r, CellsDev c, float2* n0, float2* n1, float2* n2, float2* n3, float2* n4, const float2* c0,
                           const float2* c1, const float2* c2, const float2* c3, const float2* c4) {
    int n = r.cell_end - r.cell_begin;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        uint8_t f = c.flags[r.cell_begin + k];
        if ((f & 1u) && (f & 4u)) {
            n0[k] = c0[k];
            n1[k] = c1[k];
            n2[k] = c2[k];
            n3[k] = c3[k];
            n4[k] = c4[k];
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------------------------

// ---------------------------------------------------------------------------------------------
// The differentiable objective (no reference counterpart: the reference never evaluates it):
// f = sum_e w_e WA_e + lambda * 1/2 sum_b rho_b phi_b over one region's cells at continuous
// centres xy.  WA uses the exact exponential; pins of the region's cells sit at xy + offset, other
// pins at their initial positions.  rho_b = (sum_i overlap_b(i) scale_i [target if macro] +
// non_place_b) / (area_b target) with each density box centred at its point (not truncated);
// phi = K rho from the region's Poisson solve.  K is symmetric, so d(1/2 rho^T K rho)/d rho = phi.
// ---------------------------------------------------------------------------------------------
constexpr double kObjFix = 65536.0;  // fixed point of the overlap accumulators: 2^-16 DBU^2

__device__ __forceinline__ double obj_weight(const RegionDev& r, const CellsDev& c, int g) {
    double s = (double)c.dscale[g];
    const uint8_t f = c.flags[g];
    if ((f & 1u) && (f & 2u)) s *= (double)r.target_density;
    return s;
}

// bins [x0, x1] x [y0, y1] that may touch the box (one bin of margin, clamped to the grid)
__device__ __forceinline__ void obj_span(const RegionDev& r, double lo, double hi, int org, double size, int cnt, int& a, int& b) {
    a = (int)floor((lo - org) / size) - 1;
    b = (int)floor((hi - org) / size) + 1;
    a = max(a, 0);
    b = min(b, cnt - 1);
}

__global__ void k_obj_scatter(RegionDev r, CellsDev c, const float2* xy, unsigned long long* acc) {
    const int n = r.cell_end - r.cell_begin;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int g = r.cell_begin + k;
        const int4 d = c.dbox[g];
        const double hw = 0.5 * (double)(d.z - d.x), hh = 0.5 * (double)(d.w - d.y);
        const float2 p = xy[k];
        const double bl = (double)p.x - hw, bu = (double)p.x + hw, cl = (double)p.y - hh, cu = (double)p.y + hh;
        const double s = obj_weight(r, c, g) * kObjFix;
        int x0, x1, y0, y1;
        obj_span(r, bl, bu, r.lx, r.size_x, r.cnt_x, x0, x1);
        obj_span(r, cl, cu, r.ly, r.size_y, r.cnt_y, y0, y1);
        for (int y = y0; y <= y1; ++y) {
            const double oy = fmin(cu, (double)r.bin_y[y + 1]) - fmax(cl, (double)r.bin_y[y]);
            if (oy <= 0.0) continue;
            for (int x = x0; x <= x1; ++x) {
                const double ox = fmin(bu, (double)r.bin_x[x + 1]) - fmax(bl, (double)r.bin_x[x]);
                if (ox <= 0.0) continue;
                atomicAdd(acc + (size_t)y * r.cnt_x + x, (unsigned long long)llrint(ox * oy * s));
            }
        }
    }
}

// rho per bin into r.density (float, the solve's input); the accumulators cleared
__global__ void k_obj_rho(RegionDev r, unsigned long long* acc) {
    const size_t M = (size_t)r.cnt_x * r.cnt_y;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < M; k += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(k % r.cnt_x), y = (int)(k / r.cnt_x);
        const double a = (double)(r.bin_x[x + 1] - r.bin_x[x]) * (double)(r.bin_y[y + 1] - r.bin_y[y]) * (double)r.target_density;
        const double v = (double)(long long)acc[k] / kObjFix + (double)r.non_place[k];
        r.density[k] = (float)(v / a);
        acc[k] = 0ull;
    }
}

// sum_b rho_b phi_b as FP64 block partials
__global__ void k_obj_energy(RegionDev r, double* partials) {
    __shared__ double sh[kBlock / 64];
    const size_t M = (size_t)r.cnt_x * r.cnt_y;
    double e = 0.0;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < M; k += (size_t)gridDim.x * blockDim.x)
        e += (double)r.density[k] * (double)r.phi[k];
    e = block_sum(e, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = e;
}

__device__ __forceinline__ float2 obj_pin(const NetsDev& n, const RegionDev& r, const float2* xy, int k) {
    const int cell = n.pcell[k];
    if (cell >= r.cell_begin && cell < r.cell_end) {
        const float2 p = xy[cell - r.cell_begin];
        const int2 o = n.poff[k];
        return make_float2(p.x + (float)o.x, p.y + (float)o.y);
    }
    const int2 q = n.pxy0[k];
    return make_float2((float)q.x, (float)q.y);
}

// One net per thread: the weighted WA value (FP64 block partials); with GRAD each pin's gradient
// term dWA/dx (the reference's term expressions with the exact exponential, times the weight) at
// its cell-pin index.
template <bool GRAD>
__global__ void k_obj_wa(NetsDev n, RegionDev r, const float2* xy, float cf, double* partials) {
    __shared__ double sh[kBlock / 64];
    double acc = 0.0;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n.num_nets; e += gridDim.x * blockDim.x) {
        const int s = n.start[e], t = n.start[e + 1];
        if (t <= s) continue;
        float lx = FLT_MAX, ly = FLT_MAX, ux = -FLT_MAX, uy = -FLT_MAX;
        for (int k = s; k < t; ++k) {
            const float2 p = obj_pin(n, r, xy, k);
            lx = fminf(lx, p.x);
            ux = fmaxf(ux, p.x);
            ly = fminf(ly, p.y);
            uy = fmaxf(uy, p.y);
        }
        float a0 = 0.f, a1 = 0.f, b0 = 0.f, b1 = 0.f, c0 = 0.f, c1 = 0.f, d0 = 0.f, d1 = 0.f;
        for (int k = s; k < t; ++k) {
            const float2 p = obj_pin(n, r, xy, k);
            const float ea = expf(cf * (p.x - ux)), eb = expf(cf * (lx - p.x));
            const float ec = expf(cf * (p.y - uy)), ed = expf(cf * (ly - p.y));
            a0 += ea;
            a1 += p.x * ea;
            b0 += eb;
            b1 += p.x * eb;
            c0 += ec;
            c1 += p.y * ec;
            d0 += ed;
            d1 += p.y * ed;
        }
        const float w = n.weight[e];
        acc += (double)w * (((double)a1 / a0 - (double)b1 / b0) + ((double)c1 / c0 - (double)d1 / d0));
        if (GRAD) {
            for (int k = s; k < t; ++k) {
                const int q = n.cpi[k];
                if (q < 0) continue;
                const float2 p = obj_pin(n, r, xy, k);
                const float ea = expf(cf * (p.x - ux)), eb = expf(cf * (lx - p.x));
                const float ec = expf(cf * (p.y - uy)), ed = expf(cf * (ly - p.y));
                float gx = wa_term_max(a0, a1, ea, cf, p.x) - wa_term_min(b0, b1, eb, cf, p.x);
                float gy = wa_term_max(c0, c1, ec, cf, p.y) - wa_term_min(d0, d1, ed, cf, p.y);
                n.pgrad[q] = make_float2(gx * w, gy * w);
            }
        }
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

// Per cell: the WA gradient (its pins' terms) plus lambda times the density adjoint
// sum_b phi_b d rho_b / dx, with d overlap / dx = (height of the overlap) x ([right edge in the
// bin] - [left edge in the bin]).
__global__ void k_obj_cell_grad(RegionDev r, CellsDev c, NetsDev nt, const float2* xy, float lambda, float2* grad) {
    const int n = r.cell_end - r.cell_begin;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int g = r.cell_begin + k;
        float wx = 0.0f, wy = 0.0f;
        for (int q = c.pin_start[g]; q < c.pin_start[g + 1]; ++q) {
            const float2 t = nt.pgrad[q];
            wx += t.x;
            wy += t.y;
        }
        const int4 d = c.dbox[g];
        const double hw = 0.5 * (double)(d.z - d.x), hh = 0.5 * (double)(d.w - d.y);
        const float2 p = xy[k];
        const double bl = (double)p.x - hw, bu = (double)p.x + hw, cl = (double)p.y - hh, cu = (double)p.y + hh;
        const double s = obj_weight(r, c, g);
        int x0, x1, y0, y1;
        obj_span(r, bl, bu, r.lx, r.size_x, r.cnt_x, x0, x1);
        obj_span(r, cl, cu, r.ly, r.size_y, r.cnt_y, y0, y1);
        double gx = 0.0, gy = 0.0;
        for (int y = y0; y <= y1; ++y) {
            const double by0 = r.bin_y[y], by1 = r.bin_y[y + 1];
            const double oy = fmin(cu, by1) - fmax(cl, by0);
            if (oy <= 0.0) continue;
            const double dy = (cu > by0 && cu < by1 ? 1.0 : 0.0) - (cl > by0 && cl < by1 ? 1.0 : 0.0);
            for (int x = x0; x <= x1; ++x) {
                const double bx0 = r.bin_x[x], bx1 = r.bin_x[x + 1];
                const double ox = fmin(bu, bx1) - fmax(bl, bx0);
                if (ox <= 0.0) continue;
                const double dx = (bu > bx0 && bu < bx1 ? 1.0 : 0.0) - (bl > bx0 && bl < bx1 ? 1.0 : 0.0);
                const size_t kb = (size_t)y * r.cnt_x + x;
                const double f = (double)r.phi[kb] * s / ((bx1 - bx0) * (by1 - by0) * (double)r.target_density);
                gx += f * oy * dx;
                gy += f * ox * dy;
            }
        }
        grad[k] = make_float2(wx + lambda * (float)gx, wy + lambda * (float)gy);
    }
}

// out[q] = sum of partials[q * stride + i], i < n, in order (one block)
__global__ void k_sum_partials(const double* partials, int n, int stride, int nq, double* out) {
    __shared__ double sh[kBlock / 64];
    for (int q = 0; q < nq; ++q) {
        double v = 0.0;
        for (int i = threadIdx.x; i < n; i += blockDim.x) v += partials[(size_t)q * stride + i];
        v = block_sum(v, sh);
        if (threadIdx.x == 0) out[q] = v;
        __syncthreads();
    }
}

__global__ void k_set_u64(unsigned long long* p, unsigned long long v) { *p = v; }

static int grid_for(size_t n) {
    size_t b = (n + kBlock - 1) / kBlock;
    if (b < 1) b = 1;
    if (b > (size_t)kMaxBlocks) b = kMaxBlocks;
    return round_xcds((int)b);
}

// ---- the trial recorder (set_recorder): the grouped kernels' argument records --------------------
static thread_local TrialRecorder* t_rec = nullptr;
void set_recorder(TrialRecorder* r) { t_rec = r; }

// record kinds: the kernel and its template variant
enum RecKind {
    RK_UPDATE = 1,
    RK_WA_TILES = 2,
    RK_WA_BIG = 3,
    RK_GRADIENT = 4,
    RK_FINISH = 5,
    RK_SCATTER = 16,  // + 1 for 32-bit LDS sums, + 2 for 256-cell chunks
    RK_ROWS = 64,     // + 4 mode + (0, 1, 2: L of 256, 512, 1024; 3: any L), the tables in LDS
};

struct UpdateArgsG {
    RegionDev r;
    CellsDev c;
    PlaceArgs pa;
};
struct WaArgsG {
    NetsDev n;
    CellsDev c;
    float cx, cy;
    unsigned long long* hpwl;
    int write_rec;
};
struct GradArgsG {
    RegionDev r;
    CellsDev c;
    NetsDev nt;
    float cx, cy, penalty;
    float2* wl_out;
    float2* den_out;
    float2* sum_out;
    const float2* a_slp;
    const float2* b_slp;
    const float2* a_sum;
    double* partials;
    int nan_inject;
};
struct FinishArgsG {
    const double* partials;
    int nb_bins, nb_cells;
    Sums* out;
    const unsigned long long* hpwl;
    Sums* host_out;
    const Sums* dev_tail;
    Sums* host_tail;
    unsigned long long* dev_seq;
    unsigned long long* host_seq;
    const int* skip;
};

// true when recording (the launch is recorded, not made)
template <typename A>
static bool rec_launch(int kind, int grid, int block, unsigned lds, const A& a) {
    static_assert(sizeof(A) <= (size_t)kRecArgBytes, "a grouped kernel's arguments fit a record");
    if (!t_rec) return false;
    LaunchRec q;
    q.kind = kind;
    q.grid = grid;
    q.block = block;
    q.lds = lds;
    q.bytes = (int)sizeof(A);
    std::memcpy(q.args, &a, sizeof(A));
    t_rec->recs.push_back(q);
    return true;
}
// true when recording: a launch with no grouped form, so the record fails (nothing is launched)
static bool rec_refuse() {
    if (!t_rec) return false;
    t_rec->failed = true;
    return true;
}

int block_partials_bins(const RegionDev& r) {
    const int rpb = rows_of_block(r.cnt_x);
    return (r.cnt_y + rpb - 1) / rpb;
}
// The grid of the per-cell kernels that leave FP64 partials (the gradient, the distances): no more
// blocks than the device holds at once at the gradient's occupancy, so that no second, part-filled
// round of blocks follows the first.  Fixed per device (and so is the partials' summation order).
// (computed once per process, thread-safely: contexts may run from several host threads)
static int cells_grid(size_t n) {
    static const int cap = [] {
        int dev = 0, cus = 0, per = 0;
        if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, reinterpret_cast<const void*>(k_gradient), kBlock, 0) == hipSuccess &&
            cus > 0 && per > 0)
            return std::max(kXcds, std::min(kMaxBlocks, cus * per / kXcds * kXcds));
        return kMaxBlocks;
    }();
    return std::min(grid_for(n), cap);
}
int block_partials_cells(const RegionDev& r) { return cells_grid((size_t)(r.cell_end - r.cell_begin)); }

void launch_zero_bins(hipStream_t s, const RegionDev& r) {
    if (rec_refuse()) return;
    size_t M = (size_t)r.cnt_x * r.cnt_y;
    hipLaunchKernelGGL(k_zero_bins, dim3(grid_for(M)), dim3(kBlock), 0, s, r.inst, r.filler, M);
}

template <int MODE, int CH>
static void place_scatter_ch(hipStream_t s, const RegionDev& r, const CellsDev& c, const PlaceArgs& pa, int grid) {
    if (t_rec) {
        if (MODE != PLACE_STORED) rec_refuse();
        else rec_launch(RK_SCATTER + (r.narrow ? 1 : 0) + (CH == 256 ? 2 : 0), grid, kBlock, 0, UpdateArgsG{r, c, pa});
        return;
    }
    if (r.narrow) hipLaunchKernelGGL((k_place_scatter<MODE, unsigned, CH>), dim3(grid), dim3(kBlock), 0, s, r, c, pa);
    else hipLaunchKernelGGL((k_place_scatter<MODE, unsigned long long, CH>), dim3(grid), dim3(kBlock), 0, s, r, c, pa);
}

template <int MODE>
static void place_scatter_mode(hipStream_t s, const RegionDev& r, const CellsDev& c, const PlaceArgs& pa) {
    const int n = r.cell_end - r.cell_begin;
    const int ch = r.chunk == 256 ? 256 : kPlaceChunk;
    const int chunks = (n + ch - 1) / ch;
    const int grid = round_xcds(chunks < 4096 ? chunks : 4096);
    if (ch == 256) place_scatter_ch<MODE, 256>(s, r, c, pa, grid);
    else place_scatter_ch<MODE, kPlaceChunk>(s, r, c, pa, grid);
}

static void place_scatter(hipStream_t s, int mode, const RegionDev& r, const CellsDev& c, const PlaceArgs& pa) {
    if (r.cell_end <= r.cell_begin) return;
    if (mode == PLACE_INIT) place_scatter_mode<PLACE_INIT>(s, r, c, pa);
    else if (mode == PLACE_CENTER) place_scatter_mode<PLACE_CENTER>(s, r, c, pa);
    else place_scatter_mode<PLACE_STORED>(s, r, c, pa);
}

void launch_center_init(hipStream_t s, const RegionDev& r, const CellsDev& c, float2* cur_slp, float2* prev_slp, float2* cur,
                        unsigned long long* global_adds) {
    PlaceArgs pa{};
    pa.next_slp = cur_slp;
    pa.next = prev_slp;
    pa.cur_out = cur;
    pa.global_adds = global_adds;
    place_scatter(s, PLACE_INIT, r, c, pa);
}

void launch_center_scatter(hipStream_t s, const RegionDev& r, const CellsDev& c, const float2* xy, unsigned long long* global_adds) {
    PlaceArgs pa{};
    pa.xy = xy;
    pa.global_adds = global_adds;
    place_scatter(s, PLACE_CENTER, r, c, pa);
}

static PlaceArgs update_args(const float2* cur_slp, const float2* cur_sum, const float2* cur, float2* next, float2* next_slp, float step,
                             float coeff, unsigned long long* zero) {
    PlaceArgs pa{};
    pa.zero = zero;
    pa.cur_slp = cur_slp;
    pa.cur_sum = cur_sum;
    pa.cur = cur;
    pa.next = next;
    pa.next_slp = next_slp;
    pa.s = step;
    pa.coeff = coeff;
    return pa;
}

static int update_grid(const RegionDev& r) {
    const int n = r.cell_end - r.cell_begin;
    return std::max(1, std::min((n + kBlock - 1) / kBlock, 4096));
}

void launch_update(hipStream_t s, const RegionDev& r, const CellsDev& c, const float2* cur_slp, const float2* cur_sum, const float2* cur,
                   float2* next, float2* next_slp, float step, float coeff, unsigned long long* zero) {
    if (r.cell_end <= r.cell_begin) {
        if (zero) launch_set_u64(s, zero, 0ull);
        return;
    }
    const PlaceArgs pa = update_args(cur_slp, cur_sum, cur, next, next_slp, step, coeff, zero);
    if (rec_launch(RK_UPDATE, update_grid(r), kBlock, 0, UpdateArgsG{r, c, pa})) return;
    hipLaunchKernelGGL(k_place_update, dim3(update_grid(r)), dim3(kBlock), 0, s, r, c, pa);
}

void launch_density_scatter(hipStream_t s, const RegionDev& r, const CellsDev& c, unsigned long long* global_adds) {
    PlaceArgs pa{};
    pa.global_adds = global_adds;
    place_scatter(s, PLACE_STORED, r, c, pa);
}

void launch_bin_density(hipStream_t s, const RegionDev& r, double* partials, bool clear) {
    if (rec_refuse()) return;
    hipLaunchKernelGGL(k_bin_density, dim3(block_partials_bins(r)), dim3(kBlock), 0, s, r, partials, clear ? 1 : 0);
}

__global__ void k_copy_f64(const double* a, double* b, int n) {
    if ((int)threadIdx.x < n) b[threadIdx.x] = a[threadIdx.x];
}

void launch_copy_f64(hipStream_t s, const double* a, double* b, int n) {
    if (n > 0) hipLaunchKernelGGL(k_copy_f64, dim3(1), dim3(64), 0, s, a, b, n);
}

static int ilog2(int L) {
    int k = 0;
    while ((1 << k) < L) ++k;
    return k;
}

static void transpose(hipStream_t s, const float* in, float* out, int R, int C) {
    if (rec_refuse()) return;
    dim3 g((C + 31) / 32, (R + 31) / 32), b(32, 8);
    hipLaunchKernelGGL(k_transpose, g, b, 0, s, in, out, R, C);
}

static int rows_per_block(int L) {
    const int h = L >> 1;
    const int T = h >= 4 ? h >> 2 : 1;
    return 256 / T > 0 ? 256 / T : 1;
}

static void launch_rows(hipStream_t s, int mode, RowArgs a) {
    const int h = a.L >> 1;
    a.T = h >= 4 ? h >> 2 : 1;  // L/8 threads per row (one for the tiny lengths)
    a.r3 = (h & (h - 1)) != 0 ? 1 : 0;  // h = 3 * 2^k (the grid check admits no other)
    a.logh = ilog2(a.r3 ? h / 3 : h);
    const int rpb = rows_per_block(a.L);
    const int block = a.T * rpb;
    const int grid = (a.nrows + rpb - 1) / rpb;
    const bool spectrum = mode == ROW_SPECTRUM_POISSON;
    const bool wav = mode == ROW_POISSON || mode == ROW_SPECTRUM_POISSON;
    const size_t rows_lds = ((spectrum ? 12 : 8) + (row_mode_dual(mode) ? 8 : 0)) * (size_t)a.L * rpb;
    const size_t tab_lds = sizeof(float2) * (2 * (size_t)a.L + 1) + (wav ? 4 * (size_t)a.L : 0);
    const bool tl = rows_lds + tab_lds <= 64 * 1024;
    const size_t lds = rows_lds + (tl ? tab_lds : 0);
    if (t_rec) {
        // the variant launch_rows takes below: L of 1024, 512, 256, or any (its fields) with the
        // tables in LDS
        const bool grouped = tl && (mode == ROW_BINS_DCT2 || mode == ROW_SPECTRUM_POISSON || mode == ROW_INV_FIELDS);
        const int v = a.L == 256 ? 0 : a.L == 512 ? 1 : a.L == 1024 ? 2 : 3;
        if (!grouped) rec_refuse();
        else rec_launch(RK_ROWS + 4 * mode + v, grid, block, (unsigned)lds, a);
        return;
    }
#define GPL_ROWS(M)                                                                                        \
    do {                                                                                                   \
        if (tl && a.L == 1024) hipLaunchKernelGGL((k_rows<M, true, 10>), dim3(grid), dim3(block), lds, s, a); \
        else if (tl && a.L == 512) hipLaunchKernelGGL((k_rows<M, true, 9>), dim3(grid), dim3(block), lds, s, a); \
        else if (tl && a.L == 256) hipLaunchKernelGGL((k_rows<M, true, 8>), dim3(grid), dim3(block), lds, s, a); \
        else if (tl) hipLaunchKernelGGL((k_rows<M, true, 0>), dim3(grid), dim3(block), lds, s, a);          \
        else hipLaunchKernelGGL((k_rows<M, false, 0>), dim3(grid), dim3(block), lds, s, a);                 \
    } while (0)
    switch (mode) {
    case ROW_DCT2: GPL_ROWS(ROW_DCT2); break;
    case ROW_DCT2_SCALE: GPL_ROWS(ROW_DCT2_SCALE); break;
    case ROW_POISSON: GPL_ROWS(ROW_POISSON); break;
    case ROW_INV_FIELDS: GPL_ROWS(ROW_INV_FIELDS); break;
    case ROW_SPECTRUM_POISSON: GPL_ROWS(ROW_SPECTRUM_POISSON); break;
    case ROW_BINS_DCT2: GPL_ROWS(ROW_BINS_DCT2); break;
    default: GPL_ROWS(ROW_INV_PHI); break;
    }
#undef GPL_ROWS
}

// The three-pass solve applies when every pass runs 256-thread blocks (L/8 threads per row, so
// rows of up to 2048 bins).
static bool three_pass(const RegionDev& r) { return r.cnt_x <= 2048 && r.cnt_y <= 2048; }

// Bins are y-major D[y][x].  Three passes, each writing its output transposed for the next:
// Q1 rows along x: DCT-II -> [u][y];
// Q2 rows along y (u = row): DCT-II, halving and scale (nesterov.rs:415-424) kept in LDS, the
//    Poisson divide (425-440) and the inverse transforms along y -> Ex'[y][u] (cos), Ey'[y][u]
//    (sin), phi'[y][u];
// Q3 rows along x: the field pairs (sin along x, cos along x) (and phi = cos along x).
// bins_partials: Q1 runs the bins pass fused (ROW_BINS_DCT2, its partials there) instead of
// reading r.density.
static void solve_three_pass(hipStream_t s, const RegionDev& r, const SolveBufs& sb, bool want_phi, const int* skip,
                             double* bins_partials = nullptr) {
    const int nx = r.cnt_x, ny = r.cnt_y;
    RowArgs a{};
    a.in0 = r.density;
    a.out0 = sb.work1;  // [u][y]
    a.L = nx;
    a.nrows = ny;
    a.tw = sb.tw_x;
    a.th = sb.th_x;
    a.dw = sb.dw_x;
    a.transpose_out = 1;
    a.skip = skip;
    if (bins_partials) {
        a.inst = r.inst;
        a.filler = r.filler;
        a.npf = r.non_place_f;
        a.bin_x = r.bin_x;
        a.bin_y = r.bin_y;
        a.target = r.target_density;
        a.density = r.density;
        a.partials = bins_partials;
    }
    launch_rows(s, bins_partials ? ROW_BINS_DCT2 : ROW_DCT2, a);
    a = RowArgs{};
    a.in0 = sb.work1;
    a.out0 = sb.work0;  // (Ex', Ey') pairs [y][u][2] (work0 holds 2 cnt_x cnt_y floats)
    a.out2 = sb.work3;  // phi' [y][u]
    a.interleave = 1;
    a.L = ny;
    a.nrows = nx;
    a.tw = sb.tw_y;
    a.th = sb.th_y;
    a.dw = sb.dw_y;
    a.scale = 4.0 / (double)nx / (double)ny;
    a.wrow = sb.wx;
    a.wcol = sb.wy;
    a.want_phi = want_phi ? 1 : 0;
    a.transpose_out = 1;
    a.skip = skip;
    launch_rows(s, ROW_SPECTRUM_POISSON, a);
    a = RowArgs{};
    a.in0 = sb.work0;
    a.interleave = 1;
    a.pairs = r.field;
    a.L = nx;
    a.nrows = ny;
    a.tw = sb.tw_x;
    a.th = sb.th_x;
    a.dw = sb.dw_x;
    a.skip = skip;
    launch_rows(s, ROW_INV_FIELDS, a);
    if (want_phi) {
        a = RowArgs{};
        a.in0 = sb.work3;
        a.out0 = r.phi;
        a.L = nx;
        a.nrows = ny;
        a.tw = sb.tw_x;
        a.th = sb.th_x;
        a.dw = sb.dw_x;
        a.skip = skip;
        launch_rows(s, ROW_INV_PHI, a);
    }
}

// Grids beyond the three-pass form: bins are y-major D[y][x].  P1: rows along x, DCT-II -> B[y][u]; transpose -> [u][y];
// P2: rows along y, DCT-II + halving + scale -> A[u][v]; P3: rows along v, the Poisson divide and
// the inverse transforms along y -> Ex'[u][y] (cos), Ey'[u][y] (sin), phi'[u][y]; transpose ->
// [y][u]; P4: rows along u -> the field pairs (sin along x, cos along x), phi = cos along x.
// The fused form needs the bins pass's layout in the row pass: L a power of two (so that the row
// blocks are kBlock threads).
static bool bins_fusable(const RegionDev& r) {
    return three_pass(r) && (r.cnt_x & (r.cnt_x - 1)) == 0 && row_threads(r.cnt_x) * rows_of_block(r.cnt_x) == kBlock;
}

void launch_bins_solve(hipStream_t s, const RegionDev& r, const SolveBufs& sb, double* partials, const int* skip) {
    if (!bins_fusable(r)) {
        launch_bin_density(s, r, partials);
        launch_solve(s, r, sb, false, skip);
        return;
    }
    solve_three_pass(s, r, sb, false, skip, partials);
}

void launch_solve(hipStream_t s, const RegionDev& r, const SolveBufs& sb, bool want_phi, const int* skip) {
    if (three_pass(r)) {
        solve_three_pass(s, r, sb, want_phi, skip);
        return;
    }
    const int nx = r.cnt_x, ny = r.cnt_y;
    RowArgs a{};
    // P1
    a.in0 = r.density;
    a.out0 = sb.work0;
    a.L = nx;
    a.nrows = ny;
    a.tw = sb.tw_x;
    a.th = sb.th_x;
    a.dw = sb.dw_x;
    a.skip = skip;
    launch_rows(s, ROW_DCT2, a);
    transpose(s, sb.work0, sb.work1, ny, nx);  // [u][y]
    // P2
    a = RowArgs{};
    a.in0 = sb.work1;
    a.out0 = sb.work0;
    a.L = ny;
    a.nrows = nx;
    a.tw = sb.tw_y;
    a.th = sb.th_y;
    a.dw = sb.dw_y;
    a.scale = 4.0 / (double)nx / (double)ny;
    a.skip = skip;
    launch_rows(s, ROW_DCT2_SCALE, a);  // A[u][v]
    // P3
    a = RowArgs{};
    a.in0 = sb.work0;
    a.out0 = sb.work1;  // Ex' [u][y]
    a.out1 = sb.work2;  // Ey' [u][y]
    a.out2 = sb.work3;  // phi' [u][y]
    a.L = ny;
    a.nrows = nx;
    a.tw = sb.tw_y;
    a.th = sb.th_y;
    a.dw = sb.dw_y;
    a.wrow = sb.wx;
    a.wcol = sb.wy;
    a.want_phi = want_phi ? 1 : 0;
    a.skip = skip;
    launch_rows(s, ROW_POISSON, a);
    transpose(s, sb.work1, sb.work0, nx, ny);  // Ex [y][u]
    transpose(s, sb.work2, sb.work1, nx, ny);  // Ey [y][u]
    if (want_phi) transpose(s, sb.work3, sb.work2, nx, ny);
    // P4
    a = RowArgs{};
    a.in0 = sb.work0;
    a.in1 = sb.work1;
    a.pairs = r.field;
    a.L = nx;
    a.nrows = ny;
    a.tw = sb.tw_x;
    a.th = sb.th_x;
    a.dw = sb.dw_x;
    a.skip = skip;
    launch_rows(s, ROW_INV_FIELDS, a);
    if (want_phi) {
        RowArgs p{};
        p.in0 = sb.work2;
        p.out0 = r.phi;
        p.L = nx;
        p.nrows = ny;
        p.tw = sb.tw_x;
        p.th = sb.th_x;
        p.dw = sb.dw_x;
        p.skip = skip;
        launch_rows(s, ROW_INV_PHI, p);
    }
}

template <bool SUMS>
static void wa_launch(hipStream_t s, hipStream_t s_big, const NetsDev& n, const CellsDev& c, float coef_x, float coef_y,
                      unsigned long long* hpwl, int write_rec) {
    const int nbig = n.num_nets - n.num_small;
    if (t_rec) {
        if (!SUMS) {
            rec_refuse();
            return;
        }
        const WaArgsG a{n, c, coef_x, coef_y, hpwl, write_rec};
        if (nbig > 0) rec_launch(RK_WA_BIG, nbig, kBigBlock, 0, a);
        if (n.num_small > 0) rec_launch(RK_WA_TILES, kXcds * n.tile_blocks, kBlock, 0, a);
        return;
    }
    if (nbig > 0) hipLaunchKernelGGL(k_wa_big<SUMS>, dim3(nbig), dim3(kBigBlock), 0, s_big, n, c, coef_x, coef_y, hpwl, write_rec);
    if (n.num_small > 0)
        hipLaunchKernelGGL(k_wa_tiles<SUMS>, dim3(kXcds * n.tile_blocks), dim3(kBlock), 0, s, n, c, coef_x, coef_y, hpwl, write_rec);
}

void launch_wa_forward(hipStream_t s, hipStream_t s_big, const NetsDev& n, const CellsDev& c, float coef_x, float coef_y,
                       unsigned long long* hpwl, bool write_rec) {
    wa_launch<true>(s, s_big, n, c, coef_x, coef_y, hpwl, write_rec ? 1 : 0);
}

void launch_hpwl(hipStream_t s, const NetsDev& n, const CellsDev& c, unsigned long long* hpwl) {
    wa_launch<false>(s, s, n, c, 0.0f, 0.0f, hpwl, 0);
}

void launch_gradient(hipStream_t s, const RegionDev& r, const CellsDev& c, const NetsDev& n, float coef_x, float coef_y,
                     float penalty, float2* wl, float2* den, float2* sum, const float2* a_slp, const float2* b_slp,
                     const float2* a_sum, double* partials, int nan_inject) {
    if (rec_launch(RK_GRADIENT, block_partials_cells(r), kBlock, 0,
                   GradArgsG{r, c, n, coef_x, coef_y, penalty, wl, den, sum, a_slp, b_slp, a_sum, partials, nan_inject}))
        return;
    hipLaunchKernelGGL(k_gradient, dim3(block_partials_cells(r)), dim3(kBlock), 0, s, r, c, n, coef_x, coef_y, penalty, wl, den, sum,
                       a_slp, b_slp, a_sum, partials, nan_inject);
}

void launch_finish(hipStream_t s, const double* partials, int nparts_bins, int nparts_cells, Sums* out, const unsigned long long* hpwl,
                   Sums* host_out, const Sums* dev_tail, Sums* host_tail, unsigned long long* dev_seq, unsigned long long* host_seq,
                   const int* skip, const CtlArgs& ctl) {
    if (t_rec) {
        if (ctl.g) rec_refuse();
        else
            rec_launch(RK_FINISH, 1, kFinishBlock, 0,
                       FinishArgsG{partials, nparts_bins, nparts_cells, out, hpwl, host_out, dev_tail, host_tail, dev_seq, host_seq, skip});
        return;
    }
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(kFinishBlock), 0, s, partials, nparts_bins, nparts_cells, out, hpwl, host_out, dev_tail,
                       host_tail, dev_seq, host_seq, skip, ctl);
}

void launch_distance(hipStream_t s, int n, const float2* a, const float2* b, const float2* c, const float2* d, double* partials) {
    hipLaunchKernelGGL(k_distance, dim3(cells_grid(n)), dim3(kBlock), 0, s, n, a, b, c, d, partials);
}

void launch_initial_prev(hipStream_t s, const RegionDev& r, const CellsDev& c, const float2* cur_slp, const float2* cur_sum,
                         float coef, float2* prev_slp) {
    int n = r.cell_end - r.cell_begin;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_initial_prev, dim3(grid_for(n)), dim3(kBlock), 0, s, r, c, cur_slp, cur_sum, coef, prev_slp);
}

void launch_lock_fix(hipStream_t s, const RegionDev& r, const CellsDev& c, float2* const* nv, const float2* const* cv, int nvec) {
    (void)nvec;
    int n = r.cell_end - r.cell_begin;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_lock_fix, dim3(grid_for(n)), dim3(kBlock), 0, s, r, c, nv[0], nv[1], nv[2], nv[3], nv[4], cv[0], cv[1], cv[2],
                       cv[3], cv[4]);
}

void launch_objective_density(hipStream_t s, const RegionDev& r, const SolveBufs& sb, const CellsDev& c, const float2* xy) {
    const int n = r.cell_end - r.cell_begin;
    const size_t M = (size_t)r.cnt_x * r.cnt_y;
    if (n > 0) hipLaunchKernelGGL(k_obj_scatter, dim3(grid_for(n)), dim3(kBlock), 0, s, r, c, xy, r.inst);
    hipLaunchKernelGGL(k_obj_rho, dim3(grid_for(M)), dim3(kBlock), 0, s, r, r.inst);
    launch_solve(s, r, sb, true);
}

void launch_objective_value(hipStream_t s, const RegionDev& r, const NetsDev& n, const float2* xy, float coef, double* partials,
                            double* out2) {
    const int gw = grid_for(n.num_nets), gb = grid_for((size_t)r.cnt_x * r.cnt_y);
    hipLaunchKernelGGL(k_obj_wa<false>, dim3(gw), dim3(kBlock), 0, s, n, r, xy, coef, partials);
    hipLaunchKernelGGL(k_obj_energy, dim3(gb), dim3(kBlock), 0, s, r, partials + kPartials);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kBlock), 0, s, partials, gw > gb ? gw : gb, kPartials, 2, out2);
}

void launch_objective_gradient(hipStream_t s, const RegionDev& r, const CellsDev& c, const NetsDev& n, const float2* xy, float coef,
                               float lambda, double* partials, float2* grad) {
    const int gw = grid_for(n.num_nets), nc = r.cell_end - r.cell_begin;
    hipLaunchKernelGGL(k_obj_wa<true>, dim3(gw), dim3(kBlock), 0, s, n, r, xy, coef, partials);
    if (nc > 0) hipLaunchKernelGGL(k_obj_cell_grad, dim3(grid_for(nc)), dim3(kBlock), 0, s, r, c, n, xy, lambda, grad);
}

// ---------------------------------------------------------------------------------------------
// Spatial re-sort of the cells: keys from the current bins, a stable radix sort, gathers.
// ---------------------------------------------------------------------------------------------

__device__ __forceinline__ uint32_t spread_bits(uint32_t v) {  // 11 bits -> every other bit
    v &= 0x7ffu;
    v = (v | (v << 8)) & 0x00ff00ffu;
    v = (v | (v << 4)) & 0x0f0f0f0fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

__global__ void k_resort_keys(int N, const int2* ctr, const uint8_t* flags, RegionTable t, uint32_t* keys, int32_t* idx) {
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < N; g += gridDim.x * blockDim.x) {
        int b = 0;
        while (b + 1 < t.n && g >= t.begin[b + 1]) ++b;
        const int2 m = ctr[g];
        uint32_t bx = 0, by = 0;
        if (m.x != INT_MIN) {
            double fx = (double)(m.x - t.lx[b]) / t.size_x[b], fy = (double)(m.y - t.ly[b]) / t.size_y[b];
            int ix = (int)fx, iy = (int)fy;
            ix = ix < 0 ? 0 : (ix >= t.cnt_x[b] ? t.cnt_x[b] - 1 : ix);
            iy = iy < 0 ? 0 : (iy >= t.cnt_y[b] ? t.cnt_y[b] - 1 : iy);
            bx = (uint32_t)ix;
            by = (uint32_t)iy;
        }
        bx >>= kResortShift;
        by >>= kResortShift;
        // a cell taken out of its region (kCellParked) after the region's live cells
        const uint32_t parked = (flags[g] & kCellParked) ? 1u : 0u;
        keys[g] = ((uint32_t)b << (t.morton_bits + 1)) | (parked << t.morton_bits) | spread_bits(bx) | (spread_bits(by) << 1);
        idx[g] = g;
    }
}

__global__ void k_resort_inv(int N, const int32_t* perm, int32_t* inv) {
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < N; g += gridDim.x * blockDim.x) inv[perm[g]] = g;
}

__global__ void k_resort_gather(int N, const int32_t* perm, CellsDev in, CellsDev out, int32_t* cnt) {
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < N; g += gridDim.x * blockDim.x) {
        const int o = perm[g];
        out.dbox[g] = in.dbox[o];
        out.ctr[g] = in.ctr[o];
        out.dscale[g] = in.dscale[o];
        out.flags[g] = in.flags[o];
        out.pre[g] = in.pre[o];
        out.old_of[g] = in.old_of[o];
        cnt[g] = in.pin_start[o + 1] - in.pin_start[o];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) cnt[N] = 0;
}

// The segment of a block's staged segment starts (nondecreasing, ne of them) that holds element
// d: the last one starting at or before d (an empty segment shares its start with the next one).
__device__ __forceinline__ int segment_of(const int* s_dst, int ne, int d) {
    int lo = 0, hi = ne - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s_dst[mid] <= d) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// The cells' pin lists moved into the new cell order, pin-parallel: a block takes kBlock
// consecutive new cells, stages their new and old pin starts in LDS and copies their contiguous
// new pin range with consecutive threads on consecutive pins.
__global__ void __launch_bounds__(kBlock) k_resort_pins(int N, const int32_t* perm, CellsDev in, CellsDev out, int32_t* cpi) {
    __shared__ int s_dst[kBlock + 1];
    __shared__ int s_src[kBlock];
    const int g0 = (int)blockIdx.x * kBlock;
    const int ne = min(kBlock, N - g0);
    if (ne <= 0) return;  // (uniform per block)
    if ((int)threadIdx.x < ne) {
        const int g = g0 + (int)threadIdx.x;
        s_dst[threadIdx.x] = out.pin_start[g];
        s_src[threadIdx.x] = in.pin_start[perm[g]];
    }
    if (threadIdx.x == 0) s_dst[ne] = out.pin_start[g0 + ne];
    __syncthreads();
    const int d1 = s_dst[ne];
    for (int d = s_dst[0] + (int)threadIdx.x; d < d1; d += kBlock) {
        const int i = segment_of(s_dst, ne, d);
        const int np = in.cp_npin[s_src[i] + (d - s_dst[i])];
        out.cp_npin[d] = np;
        cpi[np] = d;
    }
}

__global__ void k_gather_f2(int n, int begin, const int32_t* perm, const float2* in, float2* out) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) out[k] = in[perm[begin + k] - begin];
}

__global__ void k_remap_pcell(int P, int32_t* pcell, const int32_t* inv) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < P; k += gridDim.x * blockDim.x) {
        const int c = pcell[k];
        if (c >= 0) pcell[k] = inv[c];
    }
}

// ---- net re-sort ----------------------------------------------------------------------------
__global__ void k_net_keys(int E, NetsDev n, int num_cells, unsigned long long* keys, int32_t* idx) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) {
        const int s = n.start[e], t = n.start[e + 1];
        int fc = num_cells;  // a net on no movable cell sorts after the others
        for (int k = s; k < t; ++k) {
            const int c = n.pcell[k];
            if (c >= 0 && c < fc) fc = c;
        }
        keys[e] = ((unsigned long long)(unsigned)(fc >> 12) << 8) | (unsigned long long)(unsigned)(t - s);
        idx[e] = e;
    }
}

__global__ void k_net_deg(int E, NetsDev n, const int32_t* perm, int32_t* deg) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) {
        const int o = perm[e];
        deg[e] = n.start[o + 1] - n.start[o];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) deg[E] = 0;
}

// new small net e (old perm[e]) copies its pins to their new places; each moved pin's cell-pin
// slot learns the pin's new index
// The small nets' pins moved into the new order, pin-parallel as k_resort_pins: a block takes
// kBlock consecutive new nets and copies their (contiguous) new pin range.
__global__ void __launch_bounds__(kBlock) k_net_move(int E, NetsDev n, const int32_t* perm, const int32_t* orig, NetResortBufs b,
                                                      int32_t* cp_npin) {
    __shared__ int s_dst[kBlock + 1];
    __shared__ int s_src[kBlock];
    const int e0 = (int)blockIdx.x * kBlock;
    const int ne = min(kBlock, E - e0);
    if (ne <= 0) return;  // (uniform per block)
    if ((int)threadIdx.x < ne) {
        const int e = e0 + (int)threadIdx.x, o = perm[e];
        s_dst[threadIdx.x] = b.start[e];
        s_src[threadIdx.x] = n.start[o];
        b.weight[e] = n.weight[o];
        b.orig[e] = orig[o];
    }
    if (threadIdx.x == 0) s_dst[ne] = b.start[e0 + ne];
    __syncthreads();
    const int d1 = s_dst[ne];
    for (int d = s_dst[0] + (int)threadIdx.x; d < d1; d += kBlock) {
        const int lo = segment_of(s_dst, ne, d);
        const int src = s_src[lo] + (d - s_dst[lo]);
        const int cp = n.cpi[src];
        b.pcell[d] = n.pcell[src];
        b.poff[d] = n.poff[src];
        b.pxy0[d] = n.pxy0[src];
        b.cpi[d] = cp;
        if (cp >= 0) cp_npin[cp] = d;
    }
}

// the high-fanout nets' pins and records keep their places
__global__ void k_net_tail(int E, int num_nets, int P, NetsDev n, const int32_t* orig, NetResortBufs b) {
    const int p0 = n.start[E];
    for (int k = p0 + blockIdx.x * blockDim.x + threadIdx.x; k < P; k += gridDim.x * blockDim.x) {
        b.pcell[k] = n.pcell[k];
        b.poff[k] = n.poff[k];
        b.pxy0[k] = n.pxy0[k];
        b.cpi[k] = n.cpi[k];
    }
    for (int e = E + blockIdx.x * blockDim.x + threadIdx.x; e < num_nets; e += gridDim.x * blockDim.x) {
        b.start[e + 1] = n.start[e + 1];
        b.weight[e] = n.weight[e];
        b.orig[e] = orig[e];
    }
}

size_t net_resort_temp_bytes(int E) {
    size_t a = 0, b = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, a, (unsigned long long*)nullptr, (unsigned long long*)nullptr, (int32_t*)nullptr,
                                             (int32_t*)nullptr, E, 0, 64);
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, b, (int32_t*)nullptr, (int32_t*)nullptr, E + 1);
    return a > b ? a : b;
}

void launch_net_resort(hipStream_t s, const NetsDev& n, int P, int num_cells, const int32_t* orig, CellsDev& c, NetResortBufs& b) {
    const int E = n.num_small;
    if (E <= 0) return;
    hipLaunchKernelGGL(k_net_keys, dim3(grid_for(E)), dim3(kBlock), 0, s, E, n, num_cells, b.keys_in, b.idx_in);
    size_t bytes = b.temp_bytes;
    (void)hipcub::DeviceRadixSort::SortPairs(b.temp, bytes, b.keys_in, b.keys_out, b.idx_in, b.perm, E, 0, b.end_bit, s);
    hipLaunchKernelGGL(k_net_deg, dim3(grid_for(E)), dim3(kBlock), 0, s, E, n, b.perm, b.deg);
    bytes = b.temp_bytes;
    (void)hipcub::DeviceScan::ExclusiveSum(b.temp, bytes, b.deg, b.start, E + 1, s);
    hipLaunchKernelGGL(k_net_move, dim3((E + kBlock - 1) / kBlock), dim3(kBlock), 0, s, E, n, b.perm, orig, b, c.cp_npin);
    hipLaunchKernelGGL(k_net_tail, dim3(grid_for(P - 0)), dim3(kBlock), 0, s, E, n.num_nets, P, n, orig, b);
}

// tile[t] = the first net e with start[e] >= t kTileSpan (E where none): net e is the first of
// every tile t in (start[e-1] / kTileSpan, start[e] / kTileSpan]; tiles past the last net's start
// are empty
__global__ void k_tile_bounds(int E, const int32_t* start, int T, int32_t* tile) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) {
        const int te = start[e] / kTileSpan;
        const int tp = e == 0 ? -1 : start[e - 1] / kTileSpan;
        for (int t = tp + 1; t <= te && t < T; ++t) tile[t] = e;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        for (int t = (E > 0 ? start[E - 1] / kTileSpan + 1 : 0); t <= T; ++t) tile[t] = E;
}

void launch_tile_bounds(hipStream_t s, const NetsDev& n, int32_t* tile) {
    if (n.num_small <= 0) return;
    hipLaunchKernelGGL(k_tile_bounds, dim3(grid_for(n.num_small)), dim3(kBlock), 0, s, n.num_small, n.start, n.num_tiles, tile);
}

size_t resort_temp_bytes(int N) {
    size_t a = 0, b = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, a, (uint32_t*)nullptr, (uint32_t*)nullptr, (int32_t*)nullptr, (int32_t*)nullptr, N,
                                             0, 32);
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, b, (int32_t*)nullptr, (int32_t*)nullptr, N + 1);
    return a > b ? a : b;
}

void launch_resort_sort(hipStream_t s, int N, const int2* ctr, const uint8_t* flags, const RegionTable& t, ResortBufs& b) {
    hipLaunchKernelGGL(k_resort_keys, dim3(grid_for(N)), dim3(kBlock), 0, s, N, ctr, flags, t, b.keys_in, b.idx_in);
    size_t bytes = b.temp_bytes;
    (void)hipcub::DeviceRadixSort::SortPairs(b.temp, bytes, b.keys_in, b.keys_out, b.idx_in, b.perm, N, 0, t.end_bit, s);
    hipLaunchKernelGGL(k_resort_inv, dim3(grid_for(N)), dim3(kBlock), 0, s, N, b.perm, b.inv);
}

void launch_resort_cells(hipStream_t s, int N, const CellsDev& in, const CellsDev& out, ResortBufs& b, int32_t* cpi) {
    hipLaunchKernelGGL(k_resort_gather, dim3(grid_for(N)), dim3(kBlock), 0, s, N, b.perm, in, out, b.cnt);
    size_t bytes = b.temp_bytes;
    (void)hipcub::DeviceScan::ExclusiveSum(b.temp, bytes, b.cnt, out.pin_start, N + 1, s);
    hipLaunchKernelGGL(k_resort_pins, dim3((N + kBlock - 1) / kBlock), dim3(kBlock), 0, s, N, b.perm, in, out, cpi);
}

void launch_gather_f2(hipStream_t s, int n, int begin, const int32_t* perm, const float2* in, float2* out) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_gather_f2, dim3(grid_for(n)), dim3(kBlock), 0, s, n, begin, perm, in, out);
}

__global__ void k_gather_f2_set(int n, int begin, const int32_t* perm, GatherSet g) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int from = perm[begin + k] - begin;
#pragma unroll
        for (int q = 0; q < kGatherMax; ++q)
            if (q < g.count) g.out[q][k] = g.in[q][from];
    }
}

void launch_gather_f2_set(hipStream_t s, int n, int begin, const int32_t* perm, const GatherSet& g) {
    if (n <= 0 || g.count <= 0) return;
    hipLaunchKernelGGL(k_gather_f2_set, dim3(grid_for(n)), dim3(kBlock), 0, s, n, begin, perm, g);
}

void launch_remap_pcell(hipStream_t s, int P, int32_t* pcell, const int32_t* inv) {
    if (P <= 0) return;
    hipLaunchKernelGGL(k_remap_pcell, dim3(grid_for(P)), dim3(kBlock), 0, s, P, pcell, inv);
}

void launch_set_u64(hipStream_t s, unsigned long long* p, unsigned long long v) {
    if (rec_refuse()) return;
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, s, p, v);
}

// ---- the grouped kernels: record g of a launch takes the blocks boff[g] .. boff[g+1]-1, as block
// blockIdx.x - boff[g] of boff[g+1] - boff[g] (with per > 0: blocks g per .. (g+1) per - 1) ---------
constexpr int kGroupMax = 63;  // records per grouped launch (one wave's ballot finds a block's record)
struct GroupPos {
    unsigned g, bid, nb;
};
__device__ __forceinline__ GroupPos group_pos(unsigned per, const int* boff, int n) {
    if (per) {
        const unsigned g = blockIdx.x / per;
        return GroupPos{g, blockIdx.x - g * per, per};
    }
    const int lane = threadIdx.x & 63;
    const int v = lane <= n ? boff[lane] : INT_MAX;
    const int g = (int)__popcll(__ballot(v <= (int)blockIdx.x)) - 1;
    const int lo = __shfl(v, g, 64), hi = __shfl(v, g + 1, 64);
    return GroupPos{(unsigned)g, blockIdx.x - (unsigned)lo, (unsigned)(hi - lo)};
}

__global__ void __launch_bounds__(kBlock) k_group_update(const UpdateArgsG* t, unsigned per, const int* boff, int n) {
    const GroupPos q = group_pos(per, boff, n);
    place_update_body(t[q.g].r, t[q.g].c, t[q.g].pa, q.bid, q.nb);
}

template <typename V, int CH>
__global__ void __launch_bounds__(kBlock) k_group_scatter(const UpdateArgsG* t, unsigned per, const int* boff, int n) {
    const GroupPos q = group_pos(per, boff, n);
    place_scatter_body<PLACE_STORED, V, CH>(t[q.g].r, t[q.g].c, t[q.g].pa, q.bid, q.nb);
}

template <int MODE, int LOGL>
__global__ void __launch_bounds__(256) k_group_rows(const RowArgs* t, unsigned per, const int* boff, int n) {
    const GroupPos q = group_pos(per, boff, n);
    rows_body<MODE, true, LOGL>(t[q.g], q.bid, q.nb);
}

__global__ void __launch_bounds__(kBlock) k_group_wa_tiles(const WaArgsG* t, unsigned per, const int* boff, int n) {
    const GroupPos q = group_pos(per, boff, n);
    const WaArgsG& a = t[q.g];
    wa_tiles_body<true>(a.n, a.c, a.cx, a.cy, a.hpwl, a.write_rec, q.bid, q.nb);
}

__global__ void __launch_bounds__(kBigBlock) k_group_wa_big(const WaArgsG* t, unsigned per, const int* boff, int n) {
    const GroupPos q = group_pos(per, boff, n);
    const WaArgsG& a = t[q.g];
    wa_big_body<true>(a.n, a.c, a.cx, a.cy, a.hpwl, a.write_rec, q.bid, q.nb);
}

__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6, 8)))
k_group_gradient(const GradArgsG* t, unsigned per, const int* boff, int n) {
    const GroupPos q = group_pos(per, boff, n);
    const GradArgsG& a = t[q.g];
    gradient_body(a.r, a.c, a.nt, a.cx, a.cy, a.penalty, a.wl_out, a.den_out, a.sum_out, a.a_slp, a.b_slp, a.a_sum, a.partials, a.nan_inject,
                  q.bid, q.nb);
}

__global__ void __launch_bounds__(kFinishBlock) k_group_finish(const FinishArgsG* t) {
    const FinishArgsG& a = t[blockIdx.x];
    if (a.skip && (*a.skip & 1)) return;
    finish_body(a.partials, a.nb_bins, a.nb_cells, a.out, a.hpwl, a.host_out, a.dev_tail, a.host_tail, a.dev_seq, a.host_seq, nullptr);
}

__global__ void k_stage_copy(const uint4* src, uint4* dst, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = src[i];
}

static size_t round16(size_t b) { return (b + 15) & ~(size_t)15; }

// The order a trial's kernels run in when grouped: every region's chain (update, scatter, the
// solve's three row passes, the gradient, the finishing reduction) keeps its order, the WA kernels
// come after every update; the finishing reductions last.
static int phase_of(int k) {
    if (k == RK_UPDATE) return 0;
    if (k >= RK_SCATTER && k < RK_SCATTER + 4) return 1;
    if (k >= RK_ROWS) {
        const int mode = (k - RK_ROWS) / 4;
        return mode == ROW_BINS_DCT2 ? 2 : mode == ROW_SPECTRUM_POISSON ? 3 : 4;
    }
    if (k == RK_WA_BIG) return 5;
    if (k == RK_WA_TILES) return 6;
    if (k == RK_GRADIENT) return 7;
    return 8;
}

// The grouped launches of a set of records: per phase, the records of one kernel variant (kind,
// block, LDS) together, at most kGroupMax each; the finishing reductions grouped by their place
// among their recorder's finishes (so that each recorder's last, the one that signals the host,
// runs after its others).  Each group's arguments, then its block offsets, in the staging.
struct PlanGroup {
    int kind, block, phase;
    unsigned lds;
    unsigned per;  // every member's grid when they are equal, else 0 (the offsets table)
    int blocks;
    std::vector<const LaunchRec*> mem;
    size_t args_off, boff_off;
};

static std::vector<PlanGroup> plan_grouped(const TrialRecorder* const* recs, int n, size_t& bytes) {
    std::vector<PlanGroup> gs;
    std::vector<int> open;  // the group each (phase, kind, block, lds) or finish ordinal is filling
    for (int j = 0; j < n; ++j) {
        int fin = 0;
        for (const LaunchRec& q : recs[j]->recs) {
            const int ph = phase_of(q.kind);
            const int ord = ph == 8 ? fin++ : 0;
            int at = -1;
            for (int o : open) {
                const PlanGroup& g = gs[(size_t)o];
                if (g.phase == ph && g.kind == q.kind && g.block == q.block && g.lds == q.lds && (int)g.mem.size() < kGroupMax &&
                    (ph != 8 || g.args_off == (size_t)ord)) {
                    at = o;
                    break;
                }
            }
            if (at < 0) {
                PlanGroup g{};
                g.kind = q.kind;
                g.block = q.block;
                g.phase = ph;
                g.lds = q.lds;
                g.args_off = (size_t)ord;  // (the finishes' ordinal until the layout below)
                gs.push_back(g);
                at = (int)gs.size() - 1;
                open.push_back(at);
            }
            gs[(size_t)at].mem.push_back(&q);
        }
    }
    std::stable_sort(gs.begin(), gs.end(), [](const PlanGroup& a, const PlanGroup& b) {
        if (a.phase != b.phase) return a.phase < b.phase;
        return a.phase == 8 && a.args_off < b.args_off;
    });
    size_t at = 0;
    for (PlanGroup& g : gs) {
        bool uni = true;
        int tot = 0;
        for (const LaunchRec* q : g.mem) {
            uni = uni && q->grid == g.mem[0]->grid;
            tot += q->grid;
        }
        g.per = uni ? (unsigned)g.mem[0]->grid : 0u;
        g.blocks = tot;
        g.args_off = at;
        at += round16((size_t)g.mem[0]->bytes * g.mem.size());
        g.boff_off = at;
        at += round16(sizeof(int) * (g.mem.size() + 1));
    }
    bytes = at;
    return gs;
}

size_t grouped_arg_bytes(const TrialRecorder* const* recs, int n) {
    size_t b = 0;
    plan_grouped(recs, n, b);
    return b;
}

hipError_t launch_grouped(hipStream_t s0, const TrialRecorder* const* recs, int n, unsigned char* host_stage, unsigned char* dev_stage,
                          hipStream_t side, hipEvent_t ev_fork, hipEvent_t ev_join) {
    if (n <= 0) return hipSuccess;
    for (int j = 0; j < n; ++j)
        if (recs[j]->failed) return hipErrorInvalidValue;
    size_t at = 0;
    const std::vector<PlanGroup> gs = plan_grouped(recs, n, at);
    if (gs.empty()) return hipSuccess;
    for (const PlanGroup& g : gs) {
        const size_t b = (size_t)g.mem[0]->bytes;
        int* boff = reinterpret_cast<int*>(host_stage + g.boff_off);
        int o = 0;
        for (size_t m = 0; m < g.mem.size(); ++m) {
            if ((size_t)g.mem[m]->bytes != b) return hipErrorInvalidValue;
            std::memcpy(host_stage + g.args_off + m * b, g.mem[m]->args, b);
            boff[m] = o;
            o += g.mem[m]->grid;
        }
        boff[g.mem.size()] = o;
    }
    // the arguments to the device by a small kernel reading the pinned staging (in the stream's
    // kernel order: no copy engine between the trials' kernels)
    void* hdev = nullptr;
    hipError_t e = hipHostGetDevicePointer(&hdev, host_stage, 0);
    if (e != hipSuccess) return e;
    const int n16 = (int)(at / 16);
    hipLaunchKernelGGL(k_stage_copy, dim3((unsigned)std::min(64, (n16 + 255) / 256)), dim3(256), 0, s0, (const uint4*)hdev, (uint4*)dev_stage,
                       n16);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    // with a side stream the WA kernels run there from the updates on, joined before the gradients
    const bool two = side && ev_fork && ev_join;
    bool forked = false, waited = false, joined = false;
    for (const PlanGroup& g : gs) {
        const int k = g.kind;
        const bool wa = k == RK_WA_TILES || k == RK_WA_BIG;
        if (two && !forked && g.phase > 0) {
            e = hipEventRecord(ev_fork, s0);
            if (e != hipSuccess) return e;
            forked = true;
        }
        hipStream_t s = s0;
        if (two && wa) {
            if (!waited) {
                e = hipStreamWaitEvent(side, ev_fork, 0);
                if (e != hipSuccess) return e;
                waited = true;
            }
            s = side;
        }
        if (two && waited && !joined && !wa && g.phase > 6) {
            e = hipEventRecord(ev_join, side);
            if (e == hipSuccess) e = hipStreamWaitEvent(s0, ev_join, 0);
            if (e != hipSuccess) return e;
            joined = true;
        }
        const dim3 grid((unsigned)g.blocks), block((unsigned)g.block);
        const void* p = dev_stage + g.args_off;
        const int* boff = reinterpret_cast<const int*>(dev_stage + g.boff_off);
        const int nm = (int)g.mem.size();
        const unsigned per = g.per;
        if (k == RK_UPDATE) {
            hipLaunchKernelGGL(k_group_update, grid, block, 0, s, (const UpdateArgsG*)p, per, boff, nm);
        } else if (k >= RK_SCATTER && k < RK_SCATTER + 4) {
            const UpdateArgsG* t = (const UpdateArgsG*)p;
            switch (k - RK_SCATTER) {
            case 0: hipLaunchKernelGGL((k_group_scatter<unsigned long long, kPlaceChunk>), grid, block, 0, s, t, per, boff, nm); break;
            case 1: hipLaunchKernelGGL((k_group_scatter<unsigned, kPlaceChunk>), grid, block, 0, s, t, per, boff, nm); break;
            case 2: hipLaunchKernelGGL((k_group_scatter<unsigned long long, 256>), grid, block, 0, s, t, per, boff, nm); break;
            default: hipLaunchKernelGGL((k_group_scatter<unsigned, 256>), grid, block, 0, s, t, per, boff, nm); break;
            }
        } else if (k >= RK_ROWS) {
            const RowArgs* t = (const RowArgs*)p;
            const int mode = (k - RK_ROWS) / 4, v = (k - RK_ROWS) % 4;
            const unsigned lds = g.lds;
#define GPL_GROUP_ROWS(M)                                                                                        \
    do {                                                                                                         \
        if (v == 2) hipLaunchKernelGGL((k_group_rows<M, 10>), grid, block, lds, s, t, per, boff, nm);            \
        else if (v == 1) hipLaunchKernelGGL((k_group_rows<M, 9>), grid, block, lds, s, t, per, boff, nm);        \
        else if (v == 0) hipLaunchKernelGGL((k_group_rows<M, 8>), grid, block, lds, s, t, per, boff, nm);        \
        else hipLaunchKernelGGL((k_group_rows<M, 0>), grid, block, lds, s, t, per, boff, nm);                    \
    } while (0)
            if (mode == ROW_BINS_DCT2) GPL_GROUP_ROWS(ROW_BINS_DCT2);
            else if (mode == ROW_SPECTRUM_POISSON) GPL_GROUP_ROWS(ROW_SPECTRUM_POISSON);
            else if (mode == ROW_INV_FIELDS) GPL_GROUP_ROWS(ROW_INV_FIELDS);
            else return hipErrorInvalidValue;
#undef GPL_GROUP_ROWS
        } else if (k == RK_WA_TILES) {
            hipLaunchKernelGGL(k_group_wa_tiles, grid, block, 0, s, (const WaArgsG*)p, per, boff, nm);
        } else if (k == RK_WA_BIG) {
            hipLaunchKernelGGL(k_group_wa_big, grid, block, 0, s, (const WaArgsG*)p, per, boff, nm);
        } else if (k == RK_GRADIENT) {
            hipLaunchKernelGGL(k_group_gradient, grid, block, 0, s, (const GradArgsG*)p, per, boff, nm);
        } else if (k == RK_FINISH) {
            hipLaunchKernelGGL(k_group_finish, dim3((unsigned)nm), block, 0, s, (const FinishArgsG*)p);
        } else {
            return hipErrorInvalidValue;
        }
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (two && waited && !joined) {
        e = hipEventRecord(ev_join, side);
        if (e == hipSuccess) e = hipStreamWaitEvent(s0, ev_join, 0);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// ---- the trial graph's argument-dependent kernel nodes ---------------------------------------
int find_trial_nodes(hipGraph_t g, TrialNode* out, int cap) {
    size_t n = 0;
    if (hipGraphGetNodes(g, nullptr, &n) != hipSuccess) return -1;
    std::vector<hipGraphNode_t> nodes(n);
    if (n && hipGraphGetNodes(g, nodes.data(), &n) != hipSuccess) return -1;
    const void* f_update = reinterpret_cast<const void*>(k_place_update);
    const void* f_scatters[] = {reinterpret_cast<const void*>(k_place_scatter<PLACE_STORED, unsigned, kPlaceChunk>),
                                reinterpret_cast<const void*>(k_place_scatter<PLACE_STORED, unsigned long long, kPlaceChunk>),
                                reinterpret_cast<const void*>(k_place_scatter<PLACE_STORED, unsigned, 256>),
                                reinterpret_cast<const void*>(k_place_scatter<PLACE_STORED, unsigned long long, 256>)};
    auto is_scatter = [&](const void* f) { return f == f_scatters[0] || f == f_scatters[1] || f == f_scatters[2] || f == f_scatters[3]; };
    const void* f_small = reinterpret_cast<const void*>(k_wa_tiles<true>);
    const void* f_big = reinterpret_cast<const void*>(k_wa_big<true>);
    const void* f_grad = reinterpret_cast<const void*>(k_gradient);
    int m = 0;
    for (hipGraphNode_t nd : nodes) {
        hipGraphNodeType t;
        if (hipGraphNodeGetType(nd, &t) != hipSuccess || t != hipGraphNodeTypeKernel) continue;
        hipKernelNodeParams p{};
        if (hipGraphKernelNodeGetParams(nd, &p) != hipSuccess) return -1;
        int kind = p.func == f_update                               ? TK_UPDATE
                   : is_scatter(p.func)                             ? TK_SCATTER
                   : p.func == f_small                              ? TK_WA_SMALL
                   : p.func == f_big                                ? TK_WA_BIG
                   : p.func == f_grad                               ? TK_GRADIENT
                                                                    : -1;
        if (kind < 0) continue;
        int begin = -1;
        if (kind == TK_UPDATE || kind == TK_SCATTER || kind == TK_GRADIENT) {
            if (!p.kernelParams || !p.kernelParams[0]) return -1;
            begin = static_cast<const RegionDev*>(p.kernelParams[0])->cell_begin;  // the region of the node
        }
        if (m >= cap) return -1;
        out[m].node = nd;
        out[m].p = p;
        out[m].kind = kind;
        out[m].cell_begin = begin;
        ++m;
    }
    return m;
}

hipError_t set_update_node(hipGraphExec_t e, const TrialNode& n, const RegionDev& r, const CellsDev& c, const float2* cur_slp,
                           const float2* cur_sum, const float2* cur, float2* next, float2* next_slp, float step, float coeff,
                           unsigned long long* zero) {
    PlaceArgs pa = update_args(cur_slp, cur_sum, cur, next, next_slp, step, coeff, zero);
    RegionDev rr = r;
    CellsDev cc = c;
    void* args[] = {&rr, &cc, &pa};
    hipKernelNodeParams p = n.p;
    p.kernelParams = args;
    p.extra = nullptr;
    return hipGraphExecKernelNodeSetParams(e, n.node, &p);
}

hipError_t set_scatter_node(hipGraphExec_t e, const TrialNode& n, const RegionDev& r, const CellsDev& c,
                            unsigned long long* global_adds) {
    PlaceArgs pa{};
    pa.global_adds = global_adds;
    RegionDev rr = r;
    CellsDev cc = c;
    void* args[] = {&rr, &cc, &pa};
    hipKernelNodeParams p = n.p;
    p.kernelParams = args;
    p.extra = nullptr;
    return hipGraphExecKernelNodeSetParams(e, n.node, &p);
}

hipError_t set_wa_node(hipGraphExec_t e, const TrialNode& n, const NetsDev& nd, const CellsDev& c, float cx, float cy,
                       unsigned long long* hpwl) {
    NetsDev nn = nd;
    CellsDev cc = c;
    int write_rec = 0;
    void* args[] = {&nn, &cc, &cx, &cy, &hpwl, &write_rec};
    hipKernelNodeParams p = n.p;
    p.kernelParams = args;
    p.extra = nullptr;
    return hipGraphExecKernelNodeSetParams(e, n.node, &p);
}

hipError_t set_gradient_node(hipGraphExec_t e, const TrialNode& n, const RegionDev& r, const CellsDev& c, const NetsDev& nd,
                             float cx, float cy, float penalty, float2* wl, float2* den, float2* sum, const float2* a_slp,
                             const float2* b_slp, const float2* a_sum, double* partials, int nan_inject) {
    RegionDev rr = r;
    CellsDev cc = c;
    NetsDev nn = nd;
    void* args[] = {&rr, &cc, &nn, &cx, &cy, &penalty, &wl, &den, &sum, &a_slp, &b_slp, &a_sum, &partials, &nan_inject};
    hipKernelNodeParams p = n.p;
    p.kernelParams = args;
    p.extra = nullptr;
    return hipGraphExecKernelNodeSetParams(e, n.node, &p);
}

}  // namespace gplk
