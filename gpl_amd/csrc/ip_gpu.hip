// One round of the B2B initial placement on the device (initial_place.rs:203-355 updatePinInfo and
// createSparseMatrix, bicgstab.rs:29-197 the CSR build and the solver), with the host's evaluation
// orders kept so that its results are the host round's bit for bit:
//   * each net's pin flags (the first pin at each strict extreme) by one thread, in pin order;
//   * the B2B entries and right-hand-side terms of every net written at the offsets an exclusive
//     scan of the nets' counts gives, so the device arrays hold them in the host's push order;
//   * duplicates summed (a stable radix sort by row and column keeps that order, then one thread
//     adds a run in order), the right-hand side likewise, starting from the locked cells' values;
//   * BiCGSTAB's vectors in float with the host's expressions, its sparse products one thread per
//     row in column order, its FP64 dot products in the host's fixed 4096-element chunks summed in
//     order, and its scalar recurrence on the host exactly as the host solver runs it.
// The host keeps the rounds' bookkeeping (positions written back, the log line).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "host.h"

namespace {

constexpr int kB = 256;
constexpr int kDotChunk = 4096;  // the host's dot-product chunk

struct HipErr {
    hipError_t e;
    const char* what;
};
inline void ck(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipErr{e, what};
}

int grid_of(long long n) { return (int)std::max<long long>(1, std::min<long long>((n + kB - 1) / kB, 65535LL * 4)); }

template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    void alloc(size_t m) {
        if (p) ck(hipFree(p), "free");
        p = nullptr;
        n = m;
        ck(hipMalloc(&p, sizeof(T) * std::max<size_t>(m, 1)), "malloc");
    }
    void put(const std::vector<T>& v) {
        alloc(v.size());
        if (!v.empty()) ck(hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice), "upload");
    }
    ~DBuf() {
        if (p) (void)hipFree(p);
    }
};

// per pin, bit 0 min x, 1 max x, 2 min y, 3 max y (updatePinInfo, initial_place.rs:203-250)
__global__ void k_ip_flags(int E, const int32_t* nstart, const int32_t* npins, const int2* pc, uint8_t* flags) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) {
        int lx = INT_MAX, ly = INT_MAX, ux = INT_MIN, uy = INT_MIN;
        int pminx = -1, pminy = -1, pmaxx = -1, pmaxy = -1;
        for (int q = nstart[e]; q < nstart[e + 1]; ++q) {
            const int k = npins[q];
            const int2 c = pc[k];
            uint8_t f = flags[k];
            if (lx > c.x) {
                if (pminx >= 0) flags[pminx] &= ~1u;
                lx = c.x;
                pminx = k;
                f |= 1u;
            }
            if (ux < c.x) {
                if (pmaxx >= 0) flags[pmaxx] &= ~2u;
                ux = c.x;
                pmaxx = k;
                f |= 2u;
            }
            if (ly > c.y) {
                if (pminy >= 0) flags[pminy] &= ~4u;
                ly = c.y;
                pminy = k;
                f |= 4u;
            }
            if (uy < c.y) {
                if (pmaxy >= 0) flags[pmaxy] &= ~8u;
                uy = c.y;
                pmaxy = k;
                f |= 8u;
            }
            flags[k] = f;  // the clears above touch only earlier pins of the net
        }
    }
}

struct IpArgs {
    int E, n;
    const int32_t* nstart;
    const int32_t* npins;
    const int2* pc;         // per pin: position
    const int32_t* pinst;   // per pin: placer instance, -1 a port
    const uint8_t* flags;   // per pin
    const int2* ictr;       // per placer instance: centre
    const uint8_t* ilock;   // per placer instance: 1 locked, 2 fixed
    const int32_t* ext;     // per placer instance: its row, -1
    int max_fanout, min_diff;
    float scale;
};

// createSparseMatrix's per-net loop (initial_place.rs:279-355).  WRITE = false: counts the entries
// and right-hand-side terms each net adds along x and y; true: writes them at the net's offsets in
// the order the host pushes them.
template <bool WRITE>
__global__ void k_ip_b2b(IpArgs a, int* tcnt_x, int* tcnt_y, int* rcnt_x, int* rcnt_y, const int* toff_x, const int* toff_y,
                         const int* roff_x, const int* roff_y, unsigned long long* tkey_x, float* tval_x, unsigned long long* tkey_y,
                         float* tval_y, int* rrow_x, float* rval_x, int* rrow_y, float* rval_y, int tbase_x, int tbase_y) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < a.E; e += gridDim.x * blockDim.x) {
        const int b0 = a.nstart[e], deg = a.nstart[e + 1] - b0;
        int tc[2] = {0, 0}, rc[2] = {0, 0};
        int to[2] = {0, 0}, ro[2] = {0, 0};
        if (WRITE) {
            to[0] = tbase_x + toff_x[e];
            to[1] = tbase_y + toff_y[e];
            ro[0] = roff_x[e];
            ro[1] = roff_y[e];
        }
        if (deg > 1 && deg < a.max_fanout) {
            const float net_weight = a.scale / float(deg - 1);
            for (int i1 = 1; i1 < deg; ++i1) {
                const int q1 = a.npins[b0 + i1];
                const int inst1 = a.pinst[q1];
                const uint8_t f1 = a.flags[q1];
                const int2 c1v = a.pc[q1];
                for (int i2 = 0; i2 < i1; ++i2) {
                    const int q2 = a.npins[b0 + i2];
                    const int inst2 = a.pinst[q2];
                    if (inst1 == inst2) continue;
                    const uint8_t f2 = a.flags[q2];
                    const int2 c2v = a.pc[q2];
                    const int pa = (inst1 >= 0 && !(a.ilock[inst1] & 2u)) ? inst1 : -1;
                    const int pb = (inst2 >= 0 && !(a.ilock[inst2] & 2u)) ? inst2 : -1;
                    for (int axis = 0; axis < 2; ++axis) {
                        const bool marked = axis == 0 ? ((f1 | f2) & 3u) != 0 : ((f1 | f2) & 12u) != 0;
                        if (!marked) continue;
                        const int c1 = axis == 0 ? c1v.x : c1v.y, c2 = axis == 0 ? c2v.x : c2v.y;
                        const int diff = abs(c1 - c2);
                        const float weight = diff > a.min_diff ? net_weight / float(diff) : net_weight / float(a.min_diff);
                        auto center = [&](int k) { return axis == 0 ? a.ictr[k].x : a.ictr[k].y; };
                        unsigned long long* tk = axis == 0 ? tkey_x : tkey_y;
                        float* tv = axis == 0 ? tval_x : tval_y;
                        int* rr = axis == 0 ? rrow_x : rrow_y;
                        float* rv = axis == 0 ? rval_x : rval_y;
                        auto trip = [&](int r, int c, float v) {
                            if (WRITE) {
                                tk[to[axis]] = ((unsigned long long)(unsigned)r << 32) | (unsigned)c;
                                tv[to[axis]] = v;
                                ++to[axis];
                            } else {
                                ++tc[axis];
                            }
                        };
                        auto rhs = [&](int r, float v) {
                            if (WRITE) {
                                rr[ro[axis]] = r;
                                rv[ro[axis]] = v;
                                ++ro[axis];
                            } else {
                                ++rc[axis];
                            }
                        };
                        if (pa >= 0 && pb >= 0) {
                            const bool la = a.ilock[pa] & 1u, lb = a.ilock[pb] & 1u;
                            if (!la && !lb) {
                                const int e1 = a.ext[pa], e2 = a.ext[pb];
                                trip(e1, e1, weight);
                                trip(e2, e2, weight);
                                trip(e1, e2, -weight);
                                trip(e2, e1, -weight);
                                rhs(e1, -weight * float((c1 - center(pa)) - (c2 - center(pb))));
                                rhs(e2, -weight * float((c2 - center(pb)) - (c1 - center(pa))));
                            } else if (la && !lb) {
                                const int e2 = a.ext[pb];
                                trip(e2, e2, weight);
                                rhs(e2, weight * float(c1 - (c2 - center(pb))));
                            } else if (!la && lb) {
                                const int e1 = a.ext[pa];
                                trip(e1, e1, weight);
                                rhs(e1, weight * float(c2 - (c1 - center(pa))));
                            }
                        } else if (pa < 0 && pb >= 0) {
                            const int e2 = a.ext[pb];
                            trip(e2, e2, weight);
                            rhs(e2, weight * float(c1 - (c2 - center(pb))));
                        } else if (pa >= 0 && pb < 0) {
                            const int e1 = a.ext[pa];
                            trip(e1, e1, weight);
                            rhs(e1, weight * float(c2 - (c1 - center(pa))));
                        }
                    }
                }
            }
        }
        if (!WRITE) {
            tcnt_x[e] = tc[0];
            tcnt_y[e] = tc[1];
            rcnt_x[e] = rc[0];
            rcnt_y[e] = rc[1];
        }
    }
}

__global__ void k_heads(long long m, const unsigned long long* key, int* head) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < m; i += (long long)gridDim.x * blockDim.x)
        head[i] = (i == 0 || key[i] != key[i - 1]) ? 1 : 0;
}

__global__ void k_seg_start(long long m, const int* head, const int* seg, int* start) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < m; i += (long long)gridDim.x * blockDim.x)
        if (head[i]) start[seg[i]] = (int)i;
}

// one run of equal (row, column) keys summed in the order the stable sort kept (setFromTriplets)
__global__ void k_seg_sum(int nseg, long long m, const int* start, const unsigned long long* key, const float* val, int* col, float* out,
                          int* row_cnt) {
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < nseg; s += gridDim.x * blockDim.x) {
        const int b = start[s];
        const int e = s + 1 < nseg ? start[s + 1] : (int)m;
        float v = val[b];
        for (int i = b + 1; i < e; ++i) v += val[i];
        const unsigned long long k = key[b];
        col[s] = (int)(unsigned)(k & 0xffffffffull);
        out[s] = v;
        atomicAdd(&row_cnt[(int)(k >> 32)], 1);
    }
}

// the right-hand side: each row's terms, in push order, added to its start value
__global__ void k_rhs(int n, const int* rstart, const float* rval, float* rhs) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float v = rhs[i];
        for (int q = rstart[i]; q < rstart[i + 1]; ++q) v += rval[q];
        rhs[i] = v;
    }
}

__global__ void k_count_rows(int m, const int* row, int* cnt) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) atomicAdd(&cnt[row[i]], 1);
}

__global__ void k_spmv(int n, const int* start, const int* col, const float* val, const float* x, float* y) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float t = 0.0f;
        for (int k = start[i]; k < start[i + 1]; ++k) t += val[k] * x[col[k]];
        y[i] = t;
    }
}

// the host's FP64 dot: sequential sums over 4096-element chunks, then the chunks in order.  One wave
// per chunk: the products (exact in FP64) are formed from coalesced loads into LDS, then one lane
// sums them in order.
__global__ void __launch_bounds__(64) k_dot_chunks(int n, const float* a, const float* b, double* part) {
    __shared__ double prod[kDotChunk];
    const int c = blockIdx.x;
    const int b0 = c * kDotChunk;
    const int m = min(n - b0, kDotChunk);
    for (int i = threadIdx.x; i < m; i += 64) prod[i] = double(a[b0 + i]) * double(b[b0 + i]);
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < m; ++i) s += prod[i];
        part[c] = s;
    }
}

__global__ void k_dot_finish(int nch, const double* part, double* out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;
    for (int c = 0; c < nch; ++c) s += part[c];
    *out = s;
}

__global__ void k_sub(int n, const float* a, const float* b, float* out) {  // out = a - b
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = a[i] - b[i];
}

__global__ void k_p_update(int n, const float* r, const float* v, float fb, float fw, float* p) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = r[i] + fb * (p[i] - fw * v[i]);
}

__global__ void k_s_update(int n, const float* r, const float* v, float fa, float* s) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) s[i] = r[i] - fa * v[i];
}

__global__ void k_xr_update(int n, const float* p, const float* s, const float* t, float fa, float fw, float* x, float* r) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        x[i] += fa * p[i] + fw * s[i];
        r[i] = s[i] - fw * t[i];
    }
}

__global__ void k_fill0(int n, float* x) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] = 0.0f;
}

struct DevCsr {
    int n = 0;
    DBuf<int> start, col;
    DBuf<float> val, rhs;
};

// The device twin of the host's bicgstab (design.cpp): the same scalar recurrence on the host, the
// vectors on the device.
struct DevSolver {
    int n;
    DBuf<float> r, r0, v, p, s, t, mx;
    DBuf<double> part, res;
    double* pinned = nullptr;
    explicit DevSolver(int n_) : n(n_) {
        for (DBuf<float>* b : {&r, &r0, &v, &p, &s, &t, &mx}) b->alloc(size_t(n));
        part.alloc(size_t((n + kDotChunk - 1) / kDotChunk));
        res.alloc(1);
        ck(hipHostMalloc((void**)&pinned, sizeof(double), hipHostMallocDefault), "pinned");
    }
    ~DevSolver() {
        if (pinned) (void)hipHostFree(pinned);
    }
    double dot(const float* a, const float* b) {
        const int nch = (n + kDotChunk - 1) / kDotChunk;
        if (nch == 0) return 0.0;
        hipLaunchKernelGGL(k_dot_chunks, dim3(nch), dim3(64), 0, 0, n, a, b, part.p);
        hipLaunchKernelGGL(k_dot_finish, dim3(1), dim3(64), 0, 0, nch, part.p, res.p);
        ck(hipMemcpy(pinned, res.p, sizeof(double), hipMemcpyDeviceToHost), "dot");
        return *pinned;
    }
    void spmv(const DevCsr& m, const float* x, float* y) {
        hipLaunchKernelGGL(k_spmv, dim3(grid_of(n)), dim3(kB), 0, 0, n, m.start.p, m.col.p, m.val.p, x, y);
    }
    gpl::IpSolve solve(const DevCsr& m, float* x, int32_t max_iters) {
        const double tol = 1.1920928955078125e-07;
        const float* rhs = m.rhs.p;
        const int g = grid_of(n);
        spmv(m, x, mx.p);
        hipLaunchKernelGGL(k_sub, dim3(g), dim3(kB), 0, 0, n, rhs, mx.p, r.p);
        ck(hipMemcpy(r0.p, r.p, sizeof(float) * n, hipMemcpyDeviceToDevice), "copy");
        hipLaunchKernelGGL(k_fill0, dim3(g), dim3(kB), 0, 0, n, v.p);
        hipLaunchKernelGGL(k_fill0, dim3(g), dim3(kB), 0, 0, n, p.p);
        double r0_sqnorm = dot(r0.p, r0.p), rhs_sqnorm = dot(rhs, rhs);
        if (rhs_sqnorm == 0.0) {
            hipLaunchKernelGGL(k_fill0, dim3(g), dim3(kB), 0, 0, n, x);
            return gpl::IpSolve{max_iters, float(tol), true};
        }
        double rho = 1.0, alpha = 1.0, w = 1.0;
        const double tol2 = tol * tol * rhs_sqnorm, eps2 = tol * tol;
        int32_t i = 0, restarts = 0;
        while (dot(r.p, r.p) > tol2 && i < max_iters) {
            double rho_old = rho;
            rho = dot(r0.p, r.p);
            if (std::fabs(rho) < eps2 * r0_sqnorm) {
                spmv(m, x, mx.p);
                hipLaunchKernelGGL(k_sub, dim3(g), dim3(kB), 0, 0, n, rhs, mx.p, r.p);
                ck(hipMemcpy(r0.p, r.p, sizeof(float) * n, hipMemcpyDeviceToDevice), "copy");
                r0_sqnorm = dot(r.p, r.p);
                rho = r0_sqnorm;
                if (restarts == 0) i = 0;
                ++restarts;
            }
            double beta = (rho / rho_old) * (alpha / w);
            float fb = float(beta), fw = float(w);
            hipLaunchKernelGGL(k_p_update, dim3(g), dim3(kB), 0, 0, n, r.p, v.p, fb, fw, p.p);
            spmv(m, p.p, v.p);
            alpha = rho / dot(r0.p, v.p);
            float fa = float(alpha);
            hipLaunchKernelGGL(k_s_update, dim3(g), dim3(kB), 0, 0, n, r.p, v.p, fa, s.p);
            spmv(m, s.p, t.p);
            double tmp = dot(t.p, t.p);
            w = tmp > 0.0 ? dot(t.p, s.p) / tmp : 0.0;
            fw = float(w);
            hipLaunchKernelGGL(k_xr_update, dim3(g), dim3(kB), 0, 0, n, p.p, s.p, t.p, fa, fw, x, r.p);
            ++i;
        }
        float error = float(std::sqrt(dot(r.p, r.p) / rhs_sqnorm));
        return gpl::IpSolve{i, error, error <= float(tol)};
    }
};

template <typename T>
void exclusive_scan(const int* in, int* out, int m, DBuf<uint8_t>& tmp) {
    size_t bytes = 0;
    ck(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, m), "scan");
    if (tmp.n < bytes) tmp.alloc(bytes);
    ck(hipcub::DeviceScan::ExclusiveSum(tmp.p, bytes, in, out, m), "scan");
}

int last_value(const int* a, int m) {
    int v = 0;
    if (m > 0) ck(hipMemcpy(&v, a + m - 1, sizeof(int), hipMemcpyDeviceToHost), "copy");
    return v;
}

// One axis's CSR and right-hand side from its (key, value) entries and (row, value) terms.
void build_axis(int n, long long m, DBuf<unsigned long long>& key, DBuf<float>& val, int nr, DBuf<int>& rrow, DBuf<float>& rval,
                const std::vector<float>& rhs0, DevCsr& out, DBuf<uint8_t>& tmp) {
    int end_bit = 32;
    while (end_bit < 64 && (1ull << (end_bit - 32)) <= (unsigned long long)n) ++end_bit;
    DBuf<unsigned long long> key2;
    DBuf<float> val2;
    key2.alloc(size_t(m));
    val2.alloc(size_t(m));
    size_t bytes = 0;
    ck(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, key.p, key2.p, val.p, val2.p, (int)m, 0, end_bit), "sort");
    if (tmp.n < bytes) tmp.alloc(bytes);
    ck(hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, key.p, key2.p, val.p, val2.p, (int)m, 0, end_bit), "sort");
    DBuf<int> head, seg, start, rowcnt;
    head.alloc(size_t(m) + 1);
    seg.alloc(size_t(m) + 1);
    hipLaunchKernelGGL(k_heads, dim3(grid_of(m)), dim3(kB), 0, 0, m, key2.p, head.p);
    exclusive_scan<int>(head.p, seg.p, (int)m, tmp);
    const int nseg = m > 0 ? last_value(seg.p, (int)m) + last_value(head.p, (int)m) : 0;
    start.alloc(size_t(nseg) + 1);
    hipLaunchKernelGGL(k_seg_start, dim3(grid_of(m)), dim3(kB), 0, 0, m, head.p, seg.p, start.p);
    out.n = n;
    out.col.alloc(size_t(nseg));
    out.val.alloc(size_t(nseg));
    rowcnt.alloc(size_t(n) + 1);
    ck(hipMemset(rowcnt.p, 0, sizeof(int) * (size_t(n) + 1)), "memset");
    hipLaunchKernelGGL(k_seg_sum, dim3(grid_of(nseg)), dim3(kB), 0, 0, nseg, m, start.p, key2.p, val2.p, out.col.p, out.val.p, rowcnt.p);
    out.start.alloc(size_t(n) + 1);
    exclusive_scan<int>(rowcnt.p, out.start.p, n + 1, tmp);
    // the right-hand side: its terms stably sorted by row, then each row summed from its start value
    DBuf<int> rrow2, rcnt, rstart;
    DBuf<float> rval2;
    rrow2.alloc(size_t(nr));
    rval2.alloc(size_t(nr));
    int rbits = 1;
    while (rbits < 31 && (1 << rbits) <= n) ++rbits;
    bytes = 0;
    ck(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, rrow.p, rrow2.p, rval.p, rval2.p, nr, 0, rbits), "sort");
    if (tmp.n < bytes) tmp.alloc(bytes);
    ck(hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, rrow.p, rrow2.p, rval.p, rval2.p, nr, 0, rbits), "sort");
    rcnt.alloc(size_t(n) + 1);
    ck(hipMemset(rcnt.p, 0, sizeof(int) * (size_t(n) + 1)), "memset");
    hipLaunchKernelGGL(k_count_rows, dim3(grid_of(nr)), dim3(kB), 0, 0, nr, rrow2.p, rcnt.p);
    rstart.alloc(size_t(n) + 1);
    exclusive_scan<int>(rcnt.p, rstart.p, n + 1, tmp);
    out.rhs.put(rhs0);
    hipLaunchKernelGGL(k_rhs, dim3(grid_of(n)), dim3(kB), 0, 0, n, rstart.p, rval2.p, out.rhs.p);
}

}  // namespace

namespace gpl {

int32_t ip_round_device(gpl_design* g, const std::vector<int32_t>& ext, std::vector<float>& x, std::vector<float>& y, IpSolve* ex,
                        IpSolve* ey, std::string* err) {
    try {
        Placer& pl = g->pl;
        const gpl_opts& o = g->opts;
        const int n = int(pl.place.size());
        const int E = int(pl.net_pins.size());
        const int P = int(pl.pins.size());
        std::vector<int32_t> nstart(size_t(E) + 1, 0), npins;
        for (int e = 0; e < E; ++e) {
            npins.insert(npins.end(), pl.net_pins[e].begin(), pl.net_pins[e].end());
            nstart[e + 1] = int32_t(npins.size());
        }
        std::vector<int2> pc(P);
        std::vector<int32_t> pinst(P);
        for (int k = 0; k < P; ++k) {
            pc[k] = make_int2(pl.pins[k].cx, pl.pins[k].cy);
            pinst[k] = pl.pins[k].inst;
        }
        const int NI = int(pl.insts.size());
        std::vector<int2> ictr(NI);
        std::vector<uint8_t> ilock(NI);
        for (int k = 0; k < NI; ++k) {
            ictr[k] = make_int2(pl.insts[k].cx(), pl.insts[k].cy());
            ilock[k] = uint8_t((pl.insts[k].locked ? 1 : 0) | (pl.insts[k].fixed ? 2 : 0));
        }
        // the locked cells' diagonal entries and right-hand-side values come first (the host pushes
        // them before the nets'), then each net's
        std::vector<unsigned long long> lkey;
        std::vector<float> rx0(size_t(n), 0.0f), ry0(size_t(n), 0.0f);
        for (int e = 0; e < n; ++e) {
            const Instance& i = pl.insts[pl.place[e]];
            x[e] = float(i.cx());
            y[e] = float(i.cy());
            if (i.locked) {
                lkey.push_back(((unsigned long long)(unsigned)e << 32) | (unsigned)e);
                rx0[e] = float(i.cx());
                ry0[e] = float(i.cy());
            }
        }
        DBuf<int32_t> d_nstart, d_npins, d_pinst, d_ext;
        DBuf<int2> d_pc, d_ictr;
        DBuf<uint8_t> d_flags, d_ilock, tmp;
        d_nstart.put(nstart);
        d_npins.put(npins);
        d_pinst.put(pinst);
        d_ext.put(ext);
        d_pc.put(pc);
        d_ictr.put(ictr);
        d_ilock.put(ilock);
        d_flags.alloc(size_t(P));
        ck(hipMemset(d_flags.p, 0, std::max(P, 1)), "memset");
        hipLaunchKernelGGL(k_ip_flags, dim3(grid_of(E)), dim3(kB), 0, 0, E, d_nstart.p, d_npins.p, d_pc.p, d_flags.p);
        IpArgs a{E, n, d_nstart.p, d_npins.p, d_pc.p, d_pinst.p, d_flags.p, d_ictr.p, d_ilock.p, d_ext.p, o.initial_place_max_fanout,
                 o.initial_place_min_diff_length, o.initial_place_net_weight_scale};
        DBuf<int> tcx, tcy, rcx, rcy, tox, toy, rox, roy;
        for (DBuf<int>* b : {&tcx, &tcy, &rcx, &rcy, &tox, &toy, &rox, &roy}) b->alloc(size_t(E) + 1);
        for (DBuf<int>* b : {&tcx, &tcy, &rcx, &rcy}) ck(hipMemset(b->p, 0, sizeof(int) * (size_t(E) + 1)), "memset");
        hipLaunchKernelGGL((k_ip_b2b<false>), dim3(grid_of(E)), dim3(kB), 0, 0, a, tcx.p, tcy.p, rcx.p, rcy.p, nullptr, nullptr, nullptr,
                           nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0, 0);
        exclusive_scan<int>(tcx.p, tox.p, E + 1, tmp);
        exclusive_scan<int>(tcy.p, toy.p, E + 1, tmp);
        exclusive_scan<int>(rcx.p, rox.p, E + 1, tmp);
        exclusive_scan<int>(rcy.p, roy.p, E + 1, tmp);
        const int L = int(lkey.size());
        const long long mx_ = (long long)L + last_value(tox.p, E + 1), my_ = (long long)L + last_value(toy.p, E + 1);
        const int nrx = last_value(rox.p, E + 1), nry = last_value(roy.p, E + 1);
        DBuf<unsigned long long> kx, ky;
        DBuf<float> vx, vy, rvx, rvy;
        DBuf<int> rrx, rry;
        kx.alloc(size_t(mx_));
        ky.alloc(size_t(my_));
        vx.alloc(size_t(mx_));
        vy.alloc(size_t(my_));
        rrx.alloc(size_t(nrx));
        rry.alloc(size_t(nry));
        rvx.alloc(size_t(nrx));
        rvy.alloc(size_t(nry));
        if (L > 0) {
            std::vector<float> ones(size_t(L), 1.0f);
            ck(hipMemcpy(kx.p, lkey.data(), sizeof(unsigned long long) * L, hipMemcpyHostToDevice), "upload");
            ck(hipMemcpy(ky.p, lkey.data(), sizeof(unsigned long long) * L, hipMemcpyHostToDevice), "upload");
            ck(hipMemcpy(vx.p, ones.data(), sizeof(float) * L, hipMemcpyHostToDevice), "upload");
            ck(hipMemcpy(vy.p, ones.data(), sizeof(float) * L, hipMemcpyHostToDevice), "upload");
        }
        hipLaunchKernelGGL((k_ip_b2b<true>), dim3(grid_of(E)), dim3(kB), 0, 0, a, nullptr, nullptr, nullptr, nullptr, tox.p, toy.p, rox.p,
                           roy.p, kx.p, vx.p, ky.p, vy.p, rrx.p, rvx.p, rry.p, rvy.p, L, L);
        DevCsr csx, csy;
        build_axis(n, mx_, kx, vx, nrx, rrx, rvx, rx0, csx, tmp);
        build_axis(n, my_, ky, vy, nry, rry, rvy, ry0, csy, tmp);
        DBuf<float> dx, dy;
        dx.put(x);
        dy.put(y);
        DevSolver sv(n);
        *ex = sv.solve(csx, dx.p, o.initial_place_max_solver_iter);
        *ey = sv.solve(csy, dy.p, o.initial_place_max_solver_iter);
        ck(hipMemcpy(x.data(), dx.p, sizeof(float) * n, hipMemcpyDeviceToHost), "download");
        ck(hipMemcpy(y.data(), dy.p, sizeof(float) * n, hipMemcpyDeviceToHost), "download");
        return GPL_OK;
    } catch (const HipErr& e) {
        if (err) *err = strfmt("%s: %s", e.what, hipGetErrorString(e.e));
        return GPL_ERR_HIP;
    }
}

}  // namespace gpl
