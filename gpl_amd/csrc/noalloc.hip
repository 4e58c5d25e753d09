// gpl_density_scatter_noalloc: the density map of gpl_density_scatter (the scatter of
// updateBinsGCellDensityArea and the per-bin density and overflows, nesterov.rs:2218-2263) on
// arrays the caller holds in device memory, for callers that cannot allocate (code under HIP graph
// capture, serving loops, harnesses that forbid hipMalloc).  The only device memory beyond the
// caller's inputs and outputs -- the overflow reduction's block partials and its finished sums --
// lives in one file-scope device array, `scratch`, of a fixed size; the call allocates nothing.
//
// Each GPU holds its own copy of `scratch` at its own address until the process ends.  Host code
// never uses the array's name as an address: once per device, with that device current, the
// address is taken with hipGetSymbolAddress and kept, and the kernels get that pointer.  Every call
// on a GPU shares that GPU's array: calls that overlap on one GPU (on different streams, or from
// different host threads) race on it.  Nothing in it has to start at zero: the bins pass writes
// every partial the finishing reduction reads, and the reduction writes the sums.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <mutex>

#include "../../include/gpl_hip.h"
#include "kernels.h"

namespace {

// the two overflow sums' block partials (kPartials each), then one finished Sums record
constexpr size_t kScratchDoubles = 2 * (size_t)gplk::kPartials + sizeof(gplk::Sums) / sizeof(double);
static_assert(sizeof(gplk::Sums) % sizeof(double) == 0, "the Sums record is whole doubles");
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
constexpr int32_t kMaxBinsPerSide = 4096;

std::mutex g_mu;
double* g_ptr[kMaxDevices] = {};

// The device address of this device's `scratch`, looked up once per device.
int32_t scratch_of(int dev, double** out) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            fprintf(stderr, "gpl_density_scatter_noalloc: hipGetSymbolAddress: %s\n", hipGetErrorString(e));
            return GPL_ERR_HIP;
        }
        g_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_ptr[dev];
    return GPL_OK;
}

int32_t reject(const char* why) {
    fprintf(stderr, "gpl_density_scatter_noalloc: %s\n", why);
    return GPL_ERR_ARG;
}

}  // namespace

extern "C" int32_t gpl_density_scatter_noalloc(const gpl_density_io* io) {
    if (!io) return reject("null argument");
    if (io->num_cells < 0) return reject("negative cell count");
    if (io->cnt_x < 1 || io->cnt_y < 1 || io->cnt_x > kMaxBinsPerSide || io->cnt_y > kMaxBinsPerSide)
        return reject("bin grid outside 1 .. 4096 bins per side");
    if (!(io->size_x > 0.0) || !(io->size_y > 0.0)) return reject("bin sizes must be positive");
    if (io->num_cells > 0 && (!io->dbox || !io->density_scale || !io->cell_flags)) return reject("null cell array");
    if (!io->bin_x || !io->bin_y || !io->non_place_f || !io->inst_area || !io->filler_area || !io->density || !io->overflow)
        return reject("null bin array");
    if (reinterpret_cast<uintptr_t>(io->dbox) % 16 != 0 || reinterpret_cast<uintptr_t>(io->non_place_f) % 8 != 0)
        return reject("dbox must be 16-byte aligned and non_place_f 8-byte aligned");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return reject("no current HIP device");
    double* scr = nullptr;
    const int32_t rc = scratch_of(dev, &scr);
    if (rc != GPL_OK) return rc;

    gplk::RegionDev r{};
    r.lx = io->lx;
    r.ly = io->ly;
    r.ux = io->ux;
    r.uy = io->uy;
    r.cnt_x = io->cnt_x;
    r.cnt_y = io->cnt_y;
    r.size_x = io->size_x;
    r.size_y = io->size_y;
    r.target_density = io->target_density;
    r.cell_begin = 0;
    r.cell_end = io->num_cells;
    r.narrow = 0;         // 64-bit LDS sums: correct for any bin size (the integers are the same)
    r.edges_rounded = 0;  // the bin edges are read from the caller's tables
    r.grad_stage = 0;
    r.bin_x = io->bin_x;
    r.bin_y = io->bin_y;
    r.inst = reinterpret_cast<unsigned long long*>(io->inst_area);
    r.filler = reinterpret_cast<unsigned long long*>(io->filler_area);
    r.non_place_f = reinterpret_cast<const float2*>(io->non_place_f);
    r.density = io->density;

    gplk::CellsDev c{};
    c.dbox = reinterpret_cast<int4*>(const_cast<int32_t*>(io->dbox));  // read only (stored-box mode)
    c.dscale = const_cast<float*>(io->density_scale);
    c.flags = const_cast<uint8_t*>(io->cell_flags);

    double* partials = scr;
    gplk::Sums* sums = reinterpret_cast<gplk::Sums*>(scr + 2 * (size_t)gplk::kPartials);
    hipStream_t s = static_cast<hipStream_t>(io->stream);
    gplk::launch_zero_bins(s, r);                   // a fill kernel: the outputs start at zero
    gplk::launch_density_scatter(s, r, c, nullptr);  // the density boxes as given
    gplk::launch_bin_density(s, r, partials, false); // keep the integer areas for the caller
    gplk::launch_finish(s, partials, gplk::block_partials_bins(r), 0, sums, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    gplk::launch_copy_f64(s, &sums->overflow, io->overflow, 2);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        fprintf(stderr, "gpl_density_scatter_noalloc: launch: %s\n", hipGetErrorString(e));
        return GPL_ERR_HIP;
    }
    return GPL_OK;
}
