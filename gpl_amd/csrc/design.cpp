// Host side of the placer: a seeded synthetic design generator (in place of the ODB reader,
// main.rs:44-214), the placer's construction (placer_base.rs:301-648), the placer pins
// (initial_place.rs:100-152), the B2B initial placement with BiCGSTAB (initial_place.rs:154-446,
// bicgstab.rs) and the Nesterov placer's construction (nesterov.rs:787-933, 1683-2036).
//
// Every step follows the reference's rules and evaluation order, in code of its own.  Where this
// build departs (the normal variate's generator, the BiCGSTAB sums) the function says so.
#include "host.h"
#include "ziggurat_tables.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <numeric>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace gpl {

// ------------------------------------------------------------------------------------------
// Formatting and the log
// ------------------------------------------------------------------------------------------
std::string strfmt(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    char buf[1024];
    int n = vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (n < (int)sizeof buf) return std::string(buf, n < 0 ? 0 : n);
    std::string s(n + 1, '\0');
    va_start(ap, fmt);
    vsnprintf(&s[0], n + 1, fmt, ap);
    va_end(ap);
    s.resize(n);
    return s;
}

void Log::coded(const char* level, int id, const std::string& msg) {
    // The logger prints a code at most 1000 times.
    int* count = nullptr;
    for (auto& c : counts_)
        if (c.first == id) count = &c.second;
    if (!count) {
        counts_.push_back({id, 0});
        count = &counts_.back().second;
    }
    if (++*count > 1000) return;
    lines_.push_back(strfmt("[%s GPL-%04d] %s", level, id, msg.c_str()));
    if (echo) fprintf(stderr, "%s\n", lines_.back().c_str());
}

std::string Log::joined() const {
    std::string s;
    for (auto& l : lines_) {
        s += l;
        s += '\n';
    }
    return s;
}

// ------------------------------------------------------------------------------------------
// Mersenne Twister: the standard 32-bit recurrence and tempering; seeding by the init_genrand
// recurrence (normal.rs:40-70 documents that boost's post-seed fix-up leaves the same outputs).
// ------------------------------------------------------------------------------------------
Mt19937::Mt19937(uint32_t seed) {
    x_[0] = seed;
    for (int i = 1; i < 624; ++i) x_[i] = 1812433253u * (x_[i - 1] ^ (x_[i - 1] >> 30)) + uint32_t(i);
    i_ = 624;
}

void Mt19937::twist() {
    const uint32_t A = 0x9908b0dfu, UP = 0x80000000u, LO = 0x7fffffffu;
    for (int j = 0; j < 624; ++j) {
        uint32_t y = (x_[j] & UP) | (x_[(j + 1) % 624] & LO);
        uint32_t v = x_[(j + 397) % 624] ^ (y >> 1);
        if (y & 1u) v ^= A;
        x_[j] = v;
    }
    i_ = 0;
}

uint32_t Mt19937::next() {
    if (i_ >= 624) twist();
    uint32_t z = x_[i_++];
    z ^= z >> 11;
    z ^= (z << 7) & 0x9d2c5680u;
    z ^= (z << 15) & 0xefc60000u;
    z ^= z >> 18;
    return z;
}

// boost::random::uniform_int_distribution<int>(lo, hi), one draw (normal.rs:127-147): the
// engine's 2^32 outputs split in range+1 buckets, a draw past the last bucket redrawn.
int32_t uniform_int(Mt19937& eng, int32_t lo, int32_t hi) {
    uint32_t range = uint32_t(int64_t(hi) - int64_t(lo));
    if (range == 0) return lo;
    const uint32_t brange = 0xffffffffu;
    if (range == brange) return int32_t(int64_t(lo) + int64_t(eng.next()));
    uint32_t bucket = brange / (range + 1);
    if (brange % (range + 1) == range) bucket += 1;
    for (;;) {
        uint32_t r = eng.next() / bucket;
        if (r <= range) return int32_t(int64_t(lo) + int64_t(r));
    }
}

double uniform01(Mt19937& eng) {
    uint64_t a = eng.next() >> 5, b = eng.next() >> 6;  // 27 + 26 bits
    return (double(a) * 67108864.0 + double(b)) * (1.0 / 9007199254740992.0);
}

// ------------------------------------------------------------------------------------------
// boost::random::normal_distribution<double> on mt19937, as the reference draws it
// (normal.rs:144-214, Boost 1.89): a 128-layer ziggurat of exp(-x^2/2) whose tail samples two
// exponentials from a 256-layer ziggurat of exp(-x). The layer tables (ziggurat_tables.h) are
// derived from the ziggurat's definition by gen_ziggurat.py. What must agree with the reference
// is the count of engine words per draw and the double arithmetic on them.
// ------------------------------------------------------------------------------------------
namespace {

struct ZigTables {
    double nx[129], ny[129], ex[257], ey[257];
    ZigTables() {
        std::memcpy(nx, gplz::kNormalX, sizeof nx);
        std::memcpy(ny, gplz::kNormalY, sizeof ny);
        std::memcpy(ex, gplz::kExpX, sizeof ex);
        std::memcpy(ey, gplz::kExpY, sizeof ey);
    }
};
const ZigTables kZig;

// Two engine words: the layer byte is the low 8 bits of the first; the fraction in [0, 1) is its
// high 24 bits over 2^24 plus the second's low 29 bits, all over 2^29 (normal.rs:150-160).
double layer_and_fraction(Mt19937& eng, uint32_t& layer) {
    uint32_t w = eng.next();
    layer = w & 0xffu;
    double r = double(w >> 8) * (1.0 / 16777216.0);
    w = eng.next();
    r += double(w & 0x1fffffffu);
    r *= 0.5 / 268435456.0;
    return r;
}

// One engine word over 2^32 (uniform_01 on a 32-bit engine, normal.rs:162-166).
double one_word_uniform(Mt19937& eng) { return double(eng.next()) * (1.0 / 4294967296.0); }

// Unit exponential (normal.rs:196-214): a point under layer i's rectangle is accepted at once;
// layer 0 overflows into the tail, which is the body again shifted by x[1]; otherwise one more
// word picks a height, tested against the chord, the tangent and then exp(-x) itself.
double unit_exponential(Mt19937& eng) {
    const double* X = kZig.ex;
    const double* Y = kZig.ey;
    double shift = 0.0;
    for (;;) {
        uint32_t i;
        double r = layer_and_fraction(eng, i);
        double x = r * X[i];
        if (x < X[i + 1]) return shift + x;
        if (i == 0) {
            shift += X[1];
            continue;
        }
        double t = one_word_uniform(eng);
        double y = Y[i] + t * (Y[i + 1] - Y[i]);
        double over_chord = (X[i] - X[i + 1]) * t - (X[i] - x);
        double over_tangent = y - (Y[i + 1] + (X[i + 1] - x) * Y[i + 1]);
        if (over_chord < 0.0 && (over_tangent < 0.0 || y < std::exp(-x))) return x + shift;
    }
}

// The unit normal (normal.rs:168-194): the layer byte's bit 0 is the sign, bits 1..7 the layer.
// Layer 0 beyond the rectangle draws the tail R + x with x ~ Exp(R), y ~ Exp(1) until 2y > x^2.
// In a wedge, the chord bounds the density from above where the curve is convex (x >= 1) and the
// tangent where it is concave; the other one bounds it from below.
double unit_normal(Mt19937& eng) {
    const double* X = kZig.nx;
    const double* Y = kZig.ny;
    for (;;) {
        uint32_t b;
        double r = layer_and_fraction(eng, b);
        double sign = double(int(b & 1u) * 2 - 1);
        uint32_t i = b >> 1;
        double x = r * X[i];
        if (x < X[i + 1]) return x * sign;
        if (i == 0) {
            const double R = X[1];
            for (;;) {
                double tx = unit_exponential(eng) / R;
                double ty = unit_exponential(eng) / 1.0;
                if (2.0 * ty > tx * tx) return (tx + R) * sign;
            }
        }
        double t = one_word_uniform(eng);
        double y = Y[i] + t * (Y[i + 1] - Y[i]);
        double chord = (X[i] - X[i + 1]) * t - (X[i] - x);
        double tangent = y - (Y[i] + (X[i] - x) * Y[i] * X[i]);
        double upper = X[i] >= 1.0 ? chord : tangent;
        double lower = X[i] >= 1.0 ? tangent : chord;
        if (upper < 0.0 && (lower < 0.0 || y < std::exp(-(x * x / 2.0)))) return x * sign;
    }
}

}  // namespace

double normal(Mt19937& eng, double mean, double sigma) { return unit_normal(eng) * sigma + mean; }

// ------------------------------------------------------------------------------------------
// The synthetic generator
// ------------------------------------------------------------------------------------------
namespace {

struct Shelf {
    int32_t x, y, h;
};

int32_t geom_degree(Mt19937& rng, double p, int32_t cap) {
    int32_t k = 2;
    while (k < cap && uniform01(rng) >= p) ++k;
    return k;
}

// Picks `deg` distinct cells for a net: a centre uniform in [lo, hi), the rest from a window of
// `win` indices about it (wider when the window is too small for the degree).
void pick_cells(Mt19937& rng, int32_t lo, int32_t hi, int32_t deg, int32_t win, std::vector<int32_t>& out) {
    out.clear();
    int32_t n = hi - lo;
    if (deg > n) deg = n;
    int32_t c0 = lo + uniform_int(rng, 0, n - 1);
    out.push_back(c0);
    int32_t w = std::max(win, 4 * deg);
    int32_t a = std::max(lo, c0 - w / 2), b = std::min(hi - 1, c0 + w / 2);
    if (b - a + 1 < deg) {
        a = lo;
        b = hi - 1;
    }
    int tries = 0;
    while ((int32_t)out.size() < deg) {
        int32_t c = a + uniform_int(rng, 0, b - a);
        if (++tries > 64 * deg) {
            a = lo;
            b = hi - 1;
        }
        if (std::find(out.begin(), out.end(), c) == out.end()) out.push_back(c);
    }
}

void add_pin(Design& d, std::vector<int32_t>& term_count, Mt19937& rng, DbNet& net, int32_t inst) {
    const DbInst& di = d.insts[inst];
    int32_t hw = di.bbox.dx() / 2 - 1, hh = di.bbox.dy() / 2 - 1;
    Term t;
    t.inst = inst;
    t.off_x = hw > 0 ? uniform_int(rng, -hw, hw) : 0;
    t.off_y = hh > 0 ? uniform_int(rng, -hh, hh) : 0;
    t.order = term_count[inst]++;
    net.iterms.push_back(t);
}

}  // namespace

static int32_t generate(gpl_design* g, std::string* err) {
    const gpl_gen_config& c = g->cfg;
    Design& d = g->d;
    Mt19937 rng(c.seed);
    const int32_t rh = c.row_height, sw = c.site_width;
    d.dbu = c.dbu;
    d.site_x = sw;
    d.site_y = rh;
    const int32_t ng = c.num_groups;
    // parts: 0 the top level, g+1 group g
    std::vector<int32_t> part_cells(1, c.num_cells);
    for (int32_t k = 0; k < ng; ++k) part_cells.push_back(g->group_cells[k]);
    std::vector<int32_t> part_begin(part_cells.size() + 1, 0);
    for (size_t k = 0; k < part_cells.size(); ++k) part_begin[k + 1] = part_begin[k] + part_cells[k];
    const int32_t ncells = part_begin.back();
    if (ncells <= 0) {
        *err = "no cells";
        return GPL_ERR_ARG;
    }
    std::vector<int32_t> wsites(ncells);
    std::vector<double> part_area(part_cells.size(), 0.0);
    for (size_t k = 0; k < part_cells.size(); ++k)
        for (int32_t i = part_begin[k]; i < part_begin[k + 1]; ++i) {
            wsites[i] = uniform_int(rng, c.width_min_sites, c.width_max_sites);
            part_area[k] += double(wsites[i]) * sw * rh;
        }
    // core: explicit, or square from the utilization (regions packed with slack)
    double free_area = 0.0;
    for (size_t k = 0; k < part_cells.size(); ++k) free_area += part_area[k] / c.utilization;
    double core_area = free_area / (1.0 - c.macro_area_frac) * (ng > 0 ? 1.25 : 1.0);
    int32_t rows, sites;
    if (c.core_rows > 0 && c.core_sites > 0) {
        rows = c.core_rows;
        sites = c.core_sites;
    } else {
        double side = std::sqrt(core_area);
        rows = std::max(1, (int32_t)std::lround(side / rh));
        sites = std::max(1, (int32_t)std::lround(core_area / (double(rows) * rh) / sw));
    }
    const int32_t margin = 2 * rh;
    for (int attempt = 0;; ++attempt) {
        d.core = Rect{margin, margin, margin + sites * sw, margin + rows * rh};
        d.regions.clear();
        if (ng == 0) break;
        // shelf-pack square-ish regions, largest first
        std::vector<int32_t> order(ng);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return part_area[a + 1] > part_area[b + 1]; });
        d.regions.assign(ng, Region{});
        Shelf sh{d.core.lx, d.core.ly, 0};
        bool ok = true;
        for (int32_t gi : order) {
            double a = part_area[gi + 1] / c.utilization;
            int32_t rr = std::max(1, (int32_t)std::ceil(std::sqrt(a) / rh));
            int32_t ss = std::max(1, (int32_t)std::ceil(a / (double(rr) * rh) / sw));
            int32_t w = ss * sw, h = rr * rh;
            if (sh.x + w > d.core.ux) {
                sh.y += sh.h;
                sh.x = d.core.lx;
                sh.h = 0;
            }
            if (sh.x + w > d.core.ux || sh.y + h > d.core.uy) {
                ok = false;
                break;
            }
            d.regions[gi].box = Rect{sh.x, sh.y, sh.x + w, sh.y + h};
            sh.x += w;
            sh.h = std::max(sh.h, h);
        }
        if (ok) break;
        if (attempt > 40) {
            *err = "cannot pack the region groups";
            return GPL_ERR_ARG;
        }
        rows = (int32_t)std::ceil(rows * 1.05);
        sites = (int32_t)std::ceil(sites * 1.05);
    }
    d.die = Rect{0, 0, d.core.ux + margin, d.core.uy + margin};
    d.rows.clear();
    for (int32_t r = 0; r < rows; ++r) d.rows.push_back(Rect{d.core.lx, d.core.ly + r * rh, d.core.ux, d.core.ly + (r + 1) * rh});
    // movable cells, unplaced, at the core's lower-left
    d.insts.clear();
    for (size_t k = 0; k < part_cells.size(); ++k)
        for (int32_t i = part_begin[k]; i < part_begin[k + 1]; ++i) {
            DbInst di;
            di.bbox = Rect{d.core.lx, d.core.ly, d.core.lx + wsites[i] * sw, d.core.ly + rh};
            di.group = int32_t(k) - 1;
            d.insts.push_back(di);
        }
    // fixed macros, uniformly placed without overlap, away from the regions
    if (c.macro_area_frac > 0.0) {
        double target = c.macro_area_frac * double(d.core.area()), placed = 0.0;
        std::vector<Rect> taken;
        for (auto& r : d.regions) taken.push_back(r.box);
        int misses = 0;
        while (placed < target && misses < 2000) {
            int32_t hr = uniform_int(rng, c.macro_rows_min, c.macro_rows_max);
            hr = std::min(hr, std::max(1, rows / 2));
            double aspect = 0.5 + 1.5 * uniform01(rng);
            int32_t ws = std::max(1, (int32_t)std::lround(hr * double(rh) * aspect / sw));
            ws = std::min(ws, std::max(1, sites / 2));
            int32_t w = ws * sw, h = hr * rh;
            int32_t x = d.core.lx + uniform_int(rng, 0, std::max(0, sites - ws)) * sw;
            int32_t y = d.core.ly + uniform_int(rng, 0, std::max(0, rows - hr)) * rh;
            Rect m{x, y, x + w, y + h};
            bool clash = false;
            for (auto& t : taken)
                if (m.lx < t.ux + sw && t.lx < m.ux + sw && m.ly < t.uy + rh && t.ly < m.uy + rh) {
                    clash = true;
                    break;
                }
            if (clash) {
                ++misses;
                continue;
            }
            taken.push_back(m);
            DbInst di;
            di.bbox = m;
            di.fixed = true;
            di.block = true;
            di.placed = true;
            d.insts.push_back(di);
            placed += double(m.area());
        }
    }
    // nets per part, proportional to its cells; high-fanout nets on the top level
    std::vector<int32_t> term_count(d.insts.size(), 0);
    std::vector<int32_t> pick;
    d.nets.clear();
    int64_t assigned = 0;
    for (size_t k = 0; k < part_cells.size(); ++k) {
        int32_t nk = (k + 1 == part_cells.size()) ? int32_t(c.num_nets - assigned)
                                                   : int32_t(int64_t(c.num_nets) * part_cells[k] / ncells);
        assigned += nk;
        if (part_cells[k] < 2) continue;
        for (int32_t e = 0; e < nk; ++e) {
            int32_t deg = geom_degree(rng, c.degree_p, c.degree_cap);
            pick_cells(rng, part_begin[k], part_begin[k + 1], deg, c.locality_window, pick);
            DbNet net;
            for (int32_t ci : pick) add_pin(d, term_count, rng, net, ci);
            d.nets.push_back(std::move(net));
        }
    }
    {
        int32_t lo = part_cells[0] >= 2 ? 0 : 0, hi = part_cells[0] >= 2 ? part_begin[1] : ncells;
        for (int32_t e = 0; e < c.num_hf_nets; ++e) {
            int32_t deg = uniform_int(rng, c.hf_min_pins, c.hf_max_pins);
            pick_cells(rng, lo, hi, deg, hi - lo, pick);
            DbNet net;
            for (int32_t ci : pick) add_pin(d, term_count, rng, net, ci);
            d.nets.push_back(std::move(net));
        }
    }
    // fixed IO ports evenly on the die perimeter, each on a net
    if (c.num_io > 0 && !d.nets.empty()) {
        double W = d.die.dx(), H = d.die.dy(), P = 2.0 * (W + H);
        for (int32_t k = 0; k < c.num_io; ++k) {
            double t = (k + 0.5) * P / c.num_io, x, y;
            if (t < W) {
                x = t;
                y = 0;
            } else if (t < W + H) {
                x = W;
                y = t - W;
            } else if (t < 2 * W + H) {
                x = W - (t - W - H);
                y = H;
            } else {
                x = 0;
                y = H - (t - 2 * W - H);
            }
            int32_t px = std::min(std::max((int32_t)x, 50), d.die.ux - 50);
            int32_t py = std::min(std::max((int32_t)y, 50), d.die.uy - 50);
            Port p;
            p.bbox = Rect{px - 50, py - 50, px + 50, py + 50};
            size_t net = size_t(int64_t(k) * int64_t(d.nets.size()) / c.num_io);
            d.nets[net].bterms.push_back(p);
        }
    }
    for (auto& n : d.nets)
        for (auto& t : n.iterms) d.insts[t.inst].signal_pins++;
    return GPL_OK;
}

// ------------------------------------------------------------------------------------------
// The placer's construction
// ------------------------------------------------------------------------------------------
namespace {

// placer_base.rs:252 extendSizeByScale: about the integer centre, each side x scale truncated.
int64_t extend_size_by_scale(Instance& i, double scale) {
    int64_t old = i.area();
    if (old == 0) return 0;
    int32_t cx = i.cx(), cy = i.cy();
    int32_t ndx = (int32_t)(double(i.ux - i.lx) * scale);
    int32_t ndy = (int32_t)(double(i.uy - i.ly) * scale);
    i.lx = cx - ndx / 2;
    i.ux = cx + ndx / 2;
    i.ly = cy - ndy / 2;
    i.uy = cy + ndy / 2;
    return i.area() - old;
}

// placer_base.rs:269 snapOutward.
void snap_outward(Instance& i, int32_t ox, int32_t oy, int32_t sx, int32_t sy) {
    auto down = [](int32_t v, int32_t o, int32_t s) { return ((v - o) / s) * s + o; };
    auto up = [](int32_t v, int32_t o, int32_t s) { return ((v + s - 1 - o) / s) * s + o; };
    i.lx = down(i.lx, ox, sx);
    i.ly = down(i.ly, oy, sy);
    i.ux = up(i.ux, ox, sx);
    i.uy = up(i.uy, oy, sy);
}

// placer_base.rs:410 getMinMaxIdx.
std::pair<int32_t, int32_t> min_max_idx(int32_t ll, int32_t uu, int32_t core_ll, int32_t site, int32_t mn, int32_t mx) {
    int32_t lower = (ll - core_ll) / site;
    int32_t rem = (uu - core_ll >= site) ? (uu - core_ll) % site : (uu - core_ll);
    int32_t upper = rem == 0 ? (uu - core_ll) / site : (uu - core_ll) / site + 1;
    return {std::max(mn, lower), std::min(mx, upper)};
}

// placer_base.rs:503 initInstsForUnusableSites: a site grid over the region, off-row and other
// regions' sites blocked, fixed instances' sites taken; each maximal run of blocked sites,
// extended down the rows while the whole run stays blocked, becomes one dummy instance.
std::vector<Instance> unusable_site_dummies(const Design& d, const Placer& pl, const Base& b) {
    enum : uint8_t { BLOCKED, ROW, FIXED };
    Rect rb = b.region_bbox;
    int32_t sx = pl.site_x, sy = pl.site_y;
    int64_t count_x = rb.dx() / sx, count_y = rb.dy() / sy;
    int32_t cx = int32_t(count_x), cy = int32_t(count_y);
    std::vector<uint8_t> grid(size_t(count_x * count_y), ROW);
    auto at = [&](int32_t i, int32_t j) { return size_t(int64_t(j) * count_x + i); };
    auto mark = [&](const Rect& r, uint8_t s) {
        auto xr = min_max_idx(r.lx, r.ux, rb.lx, sx, 0, cx);
        auto yr = min_max_idx(r.ly, r.uy, rb.ly, sy, 0, cy);
        for (int32_t i = xr.first; i < xr.second; ++i)
            for (int32_t j = yr.first; j < yr.second; ++j) grid[at(i, j)] = s;
    };
    if (b.group < 0) {
        std::fill(grid.begin(), grid.end(), BLOCKED);
        for (auto& r : d.rows) mark(r, ROW);
        for (auto& r : d.regions) mark(r.box, BLOCKED);
    }
    for (auto& inst : pl.insts) {
        if (!inst.fixed || inst.db < 0) continue;
        int32_t grp = d.insts[inst.db].group;
        if (grp != b.group) continue;
        mark(Rect{inst.lx, inst.ly, inst.ux, inst.uy}, FIXED);
    }
    std::vector<Instance> out;
    for (int32_t j = 0; j < cy; ++j) {
        int32_t i = 0;
        while (i < cx) {
            if (grid[at(i, j)] == BLOCKED) {
                int32_t start = i;
                while (i < cx && grid[at(i, j)] == BLOCKED) ++i;
                int32_t end_x = i, end_y = j + 1;
                while (end_y < cy) {
                    bool all = true;
                    for (int32_t x = start; x < end_x && all; ++x) all = grid[at(x, end_y)] == BLOCKED;
                    if (!all) break;
                    ++end_y;
                }
                for (int32_t y = j; y < end_y; ++y)
                    for (int32_t x = start; x < end_x; ++x) grid[at(x, y)] = FIXED;
                Instance dm;
                dm.lx = rb.lx + sx * start;
                dm.ly = rb.ly + sy * j;
                dm.ux = rb.lx + sx * end_x;
                dm.uy = rb.ly + sy * end_y;
                dm.fixed = true;
                out.push_back(dm);
            }
            ++i;  // the run's end cell is stepped past, as the reference's loop does
        }
    }
    return out;
}

}  // namespace

int32_t build_placer(gpl_design* g, std::string* err) {
    const Design& d = g->d;
    const gpl_opts& v = g->opts;
    Placer& pl = g->pl;
    pl = Placer{};
    pl.site_x = d.site_x;
    pl.site_y = d.site_y;
    g->log.info(1, "---- Initialize GPL Main Data Structures");
    g->log.info(2, strfmt("DBU: %d", d.dbu));
    const int32_t ROW_LIMIT = 6;
    for (size_t k = 0; k < d.insts.size(); ++k) {
        const DbInst& di = d.insts[k];
        Instance inst;
        inst.db = int32_t(k);
        inst.macro = di.block || di.bbox.dy() > ROW_LIMIT * pl.site_y;
        inst.lx = di.bbox.lx;
        inst.ly = di.bbox.ly;
        inst.ux = di.bbox.lx + di.bbox.dx();
        inst.uy = di.bbox.ly + di.bbox.dy();
        inst.fixed = di.fixed;
        if (!inst.fixed) {
            inst.lx -= v.pad_left * pl.site_x;
            inst.ux += v.pad_right * pl.site_x;
        }
        if (di.bbox.dy() > d.core.dy()) {
            *err = "instance height is larger than core";
            g->log.error(119, "instance height is larger than core.");
            return 119;
        }
        if (di.bbox.dx() > d.core.dx()) {
            *err = "instance width is larger than core";
            g->log.error(120, "instance width is larger than core.");
            return 120;
        }
        if (inst.fixed) snap_outward(inst, d.core.lx, d.core.ly, pl.site_x, pl.site_y);
        pl.insts.push_back(inst);
    }
    // pin-density area adjustment (placer_base.rs:349-381)
    int64_t total_signal_pins = 0, movable_area = 0;
    for (auto& i : pl.insts)
        if (!i.fixed) {
            total_signal_pins += d.insts[i.db].signal_pins;
            movable_area += i.area();
        }
    double avg_density = movable_area > 0 ? double(total_signal_pins) / double(movable_area) : 0.0;
    int64_t adjust = 0;
    if (!v.disable_pin_density_adjust)
        for (auto& i : pl.insts) {
            if (i.fixed) continue;
            int32_t pins = d.insts[i.db].signal_pins;
            if (pins > 2 && avg_density > 0.0) {
                double target_area = double(pins) / avg_density;
                double scale = std::min(std::max(std::sqrt(target_area / double(i.area())), 0.95), 1.2);
                adjust += extend_size_by_scale(i, scale);
            }
        }
    g->log.info(35, strfmt("Pin density area adjust:    %10.3f um^2", double(adjust) / (double(d.dbu) * d.dbu)));
    // bases: the top level, then each group
    for (int32_t grp = -1; grp < (int32_t)d.regions.size(); ++grp) {
        Base b;
        b.group = grp;
        if (grp < 0) {
            b.region_bbox = d.core;
        } else {
            Rect r = d.regions[grp].box;
            Rect bb{std::max(r.lx, d.core.lx), std::max(r.ly, d.core.ly), std::min(r.ux, d.core.ux), std::min(r.uy, d.core.uy)};
            if (bb.lx > bb.ux || bb.ly > bb.uy) bb = Rect{};
            b.region_bbox = bb;
        }
        b.region_area = b.region_bbox.area();
        for (size_t idx = 0; idx < pl.insts.size(); ++idx) {
            const Instance& inst = pl.insts[idx];
            if (inst.fixed) {
                if (grp >= 0) {
                    *err = "a fixed instance against a region's boundaries is not modelled";
                    return GPL_ERR_REFUSED;
                }
                Rect r{std::max(d.core.lx, inst.lx), std::max(d.core.ly, inst.ly), std::min(d.core.ux, inst.ux), std::min(d.core.uy, inst.uy)};
                int64_t ov = (r.lx < r.ux && r.ly < r.uy) ? r.area() : 0;
                if (ov == 0) continue;
                b.fixed_insts.push_back(int32_t(idx));
                b.non_place_insts_area += ov;
            } else {
                if (d.insts[inst.db].group != grp) continue;
                b.place_insts.push_back(int32_t(idx));
                b.place_insts_area += inst.area();
                if (inst.dy() > pl.site_y * 6)
                    b.macro_insts_area += inst.area();
                else
                    b.std_insts_area += inst.area();
            }
        }
        b.dummies = unusable_site_dummies(d, pl, b);
        for (auto& dm : b.dummies) b.non_place_insts_area += dm.area();
        // a region with no movable instance is not placed (run.rs's order skips it)
        if (b.place_insts.empty()) continue;
        pl.bases.push_back(std::move(b));
    }
    if (pl.bases.empty()) {
        *err = "no movable instance";
        return GPL_ERR_ARG;
    }
    // placer pins (initial_place.rs:100): each net's terminals, then its ports
    pl.inst_pins.assign(pl.insts.size(), {});
    std::vector<int32_t> order;
    for (auto& n : d.nets) {
        std::vector<int32_t> list;
        for (auto& t : n.iterms) {
            const DbInst& di = d.insts[t.inst];
            int32_t icx = di.bbox.dx() / 2, icy = di.bbox.dy() / 2;
            Pin p;
            p.inst = t.inst;  // design and placer instances share their index here
            p.off_x = t.off_x;
            p.off_y = t.off_y;
            p.cx = di.bbox.lx + icx + t.off_x;
            p.cy = di.bbox.ly + icy + t.off_y;
            list.push_back(int32_t(pl.pins.size()));
            pl.inst_pins[t.inst].push_back(int32_t(pl.pins.size()));
            order.push_back(t.order);
            pl.pins.push_back(p);
        }
        if (!v.skip_io)
            for (auto& b : n.bterms) {
                Pin p;
                p.port = true;
                p.cx = (b.bbox.lx + b.bbox.ux) / 2;
                p.cy = (b.bbox.ly + b.bbox.uy) / 2;
                list.push_back(int32_t(pl.pins.size()));
                order.push_back(0);
                pl.pins.push_back(p);
            }
        pl.net_pins.push_back(std::move(list));
    }
    for (auto& pins : pl.inst_pins)
        std::stable_sort(pins.begin(), pins.end(), [&](int32_t a, int32_t b) { return order[a] < order[b]; });
    pl.place.clear();
    for (size_t k = 0; k < pl.insts.size(); ++k)
        if (!pl.insts[k].fixed) pl.place.push_back(int32_t(k));
    pl.db_bbox.clear();
    for (auto& di : d.insts) pl.db_bbox.push_back(di.bbox);
    return GPL_OK;
}

// ------------------------------------------------------------------------------------------
// Initial placement: B2B net model solved by BiCGSTAB (initial_place.rs:154-446)
// ------------------------------------------------------------------------------------------
namespace {

// initial_place.rs:154 set_center_location with its pins.
void set_center_location(Placer& pl, int32_t k, int32_t x, int32_t y) {
    Instance& i = pl.insts[k];
    int32_t hx = (i.ux - i.lx) / 2, hy = (i.uy - i.ly) / 2;
    i.lx = x - hx;
    i.ly = y - hy;
    i.ux = x + hx;
    i.uy = y + hy;
    int32_t cx = i.cx(), cy = i.cy();
    for (int32_t pk : pl.inst_pins[k]) {
        pl.pins[pk].cx = cx + pl.pins[pk].off_x;
        pl.pins[pk].cy = cy + pl.pins[pk].off_y;
    }
}

struct Csr {
    int32_t n = 0;
    std::vector<int64_t> start;
    std::vector<int32_t> col;
    std::vector<float> val;
};

struct Triplet {
    int32_t r, c;
    float v;
};

// setFromTriplets with duplicates summed in triplet order, rows sorted by column.
Csr csr_from_triplets(int32_t n, const std::vector<Triplet>& t) {
    Csr m;
    m.n = n;
    std::vector<int64_t> cnt(size_t(n) + 1, 0);
    for (auto& x : t) cnt[x.r + 1]++;
    for (int32_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    std::vector<int32_t> tc(t.size());
    std::vector<float> tv(t.size());
    std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
    for (auto& x : t) {
        int64_t p = pos[x.r]++;
        tc[p] = x.c;
        tv[p] = x.v;
    }
    m.start.assign(size_t(n) + 1, 0);
    m.col.reserve(t.size());
    m.val.reserve(t.size());
    std::vector<int32_t> idx;
    for (int32_t r = 0; r < n; ++r) {
        idx.resize(size_t(cnt[r + 1] - cnt[r]));
        std::iota(idx.begin(), idx.end(), 0);
        int64_t b = cnt[r];
        std::stable_sort(idx.begin(), idx.end(), [&](int32_t a, int32_t c) { return tc[b + a] < tc[b + c]; });
        for (size_t q = 0; q < idx.size(); ++q) {
            int32_t c = tc[b + idx[q]];
            float v = tv[b + idx[q]];
            if (!m.col.empty() && int64_t(m.col.size()) > m.start[r] && m.col.back() == c)
                m.val.back() += v;
            else {
                m.col.push_back(c);
                m.val.push_back(v);
            }
        }
        m.start[r + 1] = int64_t(m.col.size());
    }
    return m;
}

void spmv(const Csr& m, const std::vector<float>& x, std::vector<float>& y) {
    y.resize(size_t(m.n));
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < m.n; ++i) {
        float tmp = 0.0f;
        for (int64_t k = m.start[i]; k < m.start[i + 1]; ++k) tmp += m.val[k] * x[m.col[k]];
        y[i] = tmp;
    }
}

// FP64 dot product in fixed 4096-element chunks summed in order (thread-count independent).
double dot(const std::vector<float>& a, const std::vector<float>& b) {
    size_t n = a.size(), nch = (n + 4095) / 4096;
    std::vector<double> part(nch, 0.0);
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < (long long)nch; ++c) {
        double s = 0.0;
        size_t e = std::min(n, size_t(c + 1) * 4096);
        for (size_t i = size_t(c) * 4096; i < e; ++i) s += double(a[i]) * double(b[i]);
        part[size_t(c)] = s;
    }
    double s = 0.0;
    for (double p : part) s += p;
    return s;
}

// BiCGSTAB with the identity preconditioner and the restart rule of bicgstab.rs:148-197; the
// vectors are f32 as there, the dot products FP64 (the reference sums in Eigen's packet order).
IpSolve bicgstab(const Csr& mat, const std::vector<float>& rhs, std::vector<float>& x, int32_t max_iters) {
    const double tol = 1.1920928955078125e-07;
    const size_t n = size_t(mat.n);
    std::vector<float> r(n), r0, v(n, 0.0f), p(n, 0.0f), s(n), t(n), mx;
    spmv(mat, x, mx);
    for (size_t i = 0; i < n; ++i) r[i] = rhs[i] - mx[i];
    r0 = r;
    double r0_sqnorm = dot(r0, r0), rhs_sqnorm = dot(rhs, rhs);
    if (rhs_sqnorm == 0.0) {
        std::fill(x.begin(), x.end(), 0.0f);
        return IpSolve{max_iters, float(tol), true};
    }
    double rho = 1.0, alpha = 1.0, w = 1.0;
    const double tol2 = tol * tol * rhs_sqnorm, eps2 = tol * tol;
    int32_t i = 0, restarts = 0;
    while (dot(r, r) > tol2 && i < max_iters) {
        double rho_old = rho;
        rho = dot(r0, r);
        if (std::fabs(rho) < eps2 * r0_sqnorm) {
            spmv(mat, x, mx);
            for (size_t k = 0; k < n; ++k) r[k] = rhs[k] - mx[k];
            r0 = r;
            r0_sqnorm = dot(r, r);
            rho = r0_sqnorm;
            if (restarts == 0) i = 0;
            ++restarts;
        }
        double beta = (rho / rho_old) * (alpha / w);
        float fb = float(beta), fw = float(w);
#pragma omp parallel for schedule(static)
        for (long long k = 0; k < (long long)n; ++k) p[k] = r[k] + fb * (p[k] - fw * v[k]);
        spmv(mat, p, v);
        alpha = rho / dot(r0, v);
        float fa = float(alpha);
#pragma omp parallel for schedule(static)
        for (long long k = 0; k < (long long)n; ++k) s[k] = r[k] - fa * v[k];
        spmv(mat, s, t);
        double tmp = dot(t, t);
        w = tmp > 0.0 ? dot(t, s) / tmp : 0.0;
        fw = float(w);
#pragma omp parallel for schedule(static)
        for (long long k = 0; k < (long long)n; ++k) {
            x[k] += fa * p[k] + fw * s[k];
            r[k] = s[k] - fw * t[k];
        }
        ++i;
    }
    float error = float(std::sqrt(dot(r, r) / rhs_sqnorm));
    return IpSolve{i, error, error <= float(tol)};
}

}  // namespace

int64_t db_hpwl(const gpl_design* g) {
    const Design& d = g->d;
    int64_t total = 0;
    for (auto& n : d.nets) {
        int32_t lx = INT32_MAX, ly = INT32_MAX, ux = INT32_MIN, uy = INT32_MIN;
        for (auto& t : n.iterms) {
            const Rect& b = g->pl.db_bbox[t.inst];
            lx = std::min(lx, b.lx);
            ly = std::min(ly, b.ly);
            ux = std::max(ux, b.ux);
            uy = std::max(uy, b.uy);
        }
        if (!g->opts.skip_io)
            for (auto& p : n.bterms) {
                lx = std::min(lx, p.bbox.lx);
                ly = std::min(ly, p.bbox.ly);
                ux = std::max(ux, p.bbox.ux);
                uy = std::max(uy, p.bbox.uy);
            }
        if (ux >= lx) total += int64_t(ux - lx) + int64_t(uy - ly);
    }
    return total;
}

int32_t initial_place(gpl_design* g, std::string* err) {
    const Design& d = g->d;
    Placer& pl = g->pl;
    const gpl_opts& o = g->opts;
    g->log.info(5, "---- Execute Conjugate Gradient Initial Placement.");
    // placeInstsInitialPositions (initial_place.rs:170): region centre, else core centre
    int32_t ccx = (d.core.lx + d.core.ux) / 2, ccy = (d.core.ly + d.core.uy) / 2;
    int n_region = 0, n_db = 0, n_core = 0;
    for (int32_t k : pl.place) {
        if (pl.insts[k].locked) continue;
        const DbInst& di = d.insts[pl.insts[k].db];
        if (di.group >= 0) {
            const Rect& r = d.regions[di.group].box;
            set_center_location(pl, k, r.ux - (r.ux - r.lx) / 2, r.uy - (r.uy - r.ly) / 2);
            ++n_region;
        } else if (di.placed) {
            set_center_location(pl, k, (di.bbox.lx + di.bbox.ux) / 2, (di.bbox.ly + di.bbox.uy) / 2);
            ++n_db;
        } else {
            set_center_location(pl, k, ccx, ccy);
            ++n_core;
        }
    }
    g->log.info(51, strfmt("Source of initial instance position counters:\n\tOdb location = %d\tCore center = %d\tRegion center = %d", n_db, n_core, n_region));
    const int32_t n = int32_t(pl.place.size());
    std::vector<int32_t> ext(pl.insts.size(), -1);
    for (int32_t e = 0; e < n; ++e) ext[pl.place[e]] = e;
    g->ip_rounds = 0;
    for (int32_t iter = 1; iter <= o.initial_place_max_iter; ++iter) {
        std::vector<float> x(size_t(n), 0.0f), y(size_t(n), 0.0f);
        IpSolve ex{}, ey{};
        if (g->ip_device) {
            const int32_t rc = ip_round_device(g, ext, x, y, &ex, &ey, err);
            if (rc != GPL_OK) return rc;
        } else {
            // updatePinInfo (initial_place.rs:203): the first pin at each strict extreme
            for (auto& p : pl.pins) p.min_x = p.max_x = p.min_y = p.max_y = false;
            for (auto& net : pl.net_pins) {
                int32_t lx = INT32_MAX, ly = INT32_MAX, ux = INT32_MIN, uy = INT32_MIN;
                int32_t pminx = -1, pminy = -1, pmaxx = -1, pmaxy = -1;
                for (int32_t k : net) {
                    Pin& p = pl.pins[k];
                    if (lx > p.cx) {
                        if (pminx >= 0) pl.pins[pminx].min_x = false;
                        lx = p.cx;
                        pminx = k;
                        p.min_x = true;
                    }
                    if (ux < p.cx) {
                        if (pmaxx >= 0) pl.pins[pmaxx].max_x = false;
                        ux = p.cx;
                        pmaxx = k;
                        p.max_x = true;
                    }
                    if (ly > p.cy) {
                        if (pminy >= 0) pl.pins[pminy].min_y = false;
                        ly = p.cy;
                        pminy = k;
                        p.min_y = true;
                    }
                    if (uy < p.cy) {
                        if (pmaxy >= 0) pl.pins[pmaxy].max_y = false;
                        uy = p.cy;
                        pmaxy = k;
                        p.max_y = true;
                    }
                }
            }
            // createSparseMatrix (initial_place.rs:279)
            std::vector<Triplet> tx, ty;
            std::vector<float> bx(size_t(n), 0.0f), by(size_t(n), 0.0f);
            for (int32_t e = 0; e < n; ++e) {
                const Instance& i = pl.insts[pl.place[e]];
                x[e] = float(i.cx());
                y[e] = float(i.cy());
                if (i.locked) {
                    tx.push_back({e, e, 1.0f});
                    ty.push_back({e, e, 1.0f});
                    bx[e] = float(i.cx());
                    by[e] = float(i.cy());
                }
            }
            auto placed = [&](const Pin& p) -> int32_t { return (p.inst >= 0 && !pl.insts[p.inst].fixed) ? p.inst : -1; };
            for (auto& net : pl.net_pins) {
                if (net.size() <= 1 || (int32_t)net.size() >= o.initial_place_max_fanout) continue;
                float net_weight = o.initial_place_net_weight_scale / float(net.size() - 1);
                for (size_t i1 = 1; i1 < net.size(); ++i1) {
                    const Pin& p1 = pl.pins[net[i1]];
                    for (size_t i2 = 0; i2 < i1; ++i2) {
                        const Pin& p2 = pl.pins[net[i2]];
                        if (p1.inst == p2.inst) continue;
                        for (int axis = 0; axis < 2; ++axis) {
                            bool marked = axis == 0 ? (p1.min_x || p1.max_x || p2.min_x || p2.max_x) : (p1.min_y || p1.max_y || p2.min_y || p2.max_y);
                            if (!marked) continue;
                            int32_t c1 = axis == 0 ? p1.cx : p1.cy, c2 = axis == 0 ? p2.cx : p2.cy;
                            auto& trip = axis == 0 ? tx : ty;
                            auto& rhs = axis == 0 ? bx : by;
                            auto center = [&](int32_t k) { return axis == 0 ? pl.insts[k].cx() : pl.insts[k].cy(); };
                            int32_t diff = std::abs(c1 - c2);
                            float weight = diff > o.initial_place_min_diff_length ? net_weight / float(diff) : net_weight / float(o.initial_place_min_diff_length);
                            int32_t a = placed(p1), b = placed(p2);
                            if (a >= 0 && b >= 0) {
                                bool la = pl.insts[a].locked, lb = pl.insts[b].locked;
                                if (!la && !lb) {
                                    int32_t e1 = ext[a], e2 = ext[b];
                                    trip.push_back({e1, e1, weight});
                                    trip.push_back({e2, e2, weight});
                                    trip.push_back({e1, e2, -weight});
                                    trip.push_back({e2, e1, -weight});
                                    rhs[e1] += -weight * float((c1 - center(a)) - (c2 - center(b)));
                                    rhs[e2] += -weight * float((c2 - center(b)) - (c1 - center(a)));
                                } else if (la && !lb) {
                                    int32_t e2 = ext[b];
                                    trip.push_back({e2, e2, weight});
                                    rhs[e2] += weight * float(c1 - (c2 - center(b)));
                                } else if (!la && lb) {
                                    int32_t e1 = ext[a];
                                    trip.push_back({e1, e1, weight});
                                    rhs[e1] += weight * float(c2 - (c1 - center(a)));
                                }
                            } else if (a < 0 && b >= 0) {
                                int32_t e2 = ext[b];
                                trip.push_back({e2, e2, weight});
                                rhs[e2] += weight * float(c1 - (c2 - center(b)));
                            } else if (a >= 0 && b < 0) {
                                int32_t e1 = ext[a];
                                trip.push_back({e1, e1, weight});
                                rhs[e1] += weight * float(c2 - (c1 - center(a)));
                            }
                        }
                    }
                }
            }
            Csr mx = csr_from_triplets(n, tx), my = csr_from_triplets(n, ty);
            tx.clear();
            ty.clear();
            ex = bicgstab(mx, bx, x, o.initial_place_max_solver_iter);
            ey = bicgstab(my, by, y, o.initial_place_max_solver_iter);
        }
        g->ip_rounds = iter;
        if (std::isnan(ex.error) || std::isnan(ey.error)) {
            g->log.warn(325, strfmt("Conjugate gradient initial placement solver failed at iteration %d. ", iter));
            break;
        }
        float error_max = std::max(ex.error, ey.error);
        g->ip_error = error_max;
        g->log.report(strfmt("[InitialPlace]  Iter: %d conjugate gradient residual: %.8f HPWL: %lld", iter, error_max, (long long)db_hpwl(g)));
        // updateCoordi (initial_place.rs:397): truncated, clamped to the core and the region
        for (int32_t e = 0; e < n; ++e) {
            int32_t k = pl.place[e];
            if (pl.insts[k].locked) continue;
            int32_t nx = (int32_t)x[e], ny = (int32_t)y[e];
            nx = std::min(std::max(nx, d.core.lx), d.core.ux);
            ny = std::min(std::max(ny, d.core.ly), d.core.uy);
            int32_t di = pl.insts[k].db;
            if (d.insts[di].group >= 0) {
                const Rect& r = d.regions[d.insts[di].group].box;
                nx = std::min(std::max(nx, r.lx), r.ux);
                ny = std::min(std::max(ny, r.ly), r.uy);
            }
            set_center_location(pl, k, nx, ny);
            Rect b = pl.db_bbox[di];
            pl.db_bbox[di] = Rect{pl.insts[k].lx, pl.insts[k].ly, pl.insts[k].lx + b.dx(), pl.insts[k].ly + b.dy()};
        }
        if (double(error_max) <= 1e-5 && iter >= 5) break;
    }
    return GPL_OK;
}

// ------------------------------------------------------------------------------------------
// The Nesterov placer's construction (nesterov.rs:787-933, 1683-2036)
// ------------------------------------------------------------------------------------------
namespace {

struct Cell {
    int32_t inst = -1;
    bool macro = false, locked = false;
    int32_t lx, ly, ux, uy, dlx, dly, dux, duy;
    float density_scale = 0.0f;
    int32_t dx() const { return ux - lx; }
    int32_t dy() const { return uy - ly; }
    int32_t d_cx() const { return (dux + dlx) / 2; }
    int32_t d_cy() const { return (duy + dly) / 2; }
};

Cell filler_cell(int32_t cx, int32_t cy, int32_t dx, int32_t dy) {
    Cell c;
    c.lx = c.dlx = cx - dx / 2;
    c.ly = c.dly = cy - dy / 2;
    c.ux = c.dux = cx + dx / 2;
    c.uy = c.duy = cy + dy / 2;
    return c;
}

}  // namespace

uint32_t round_down_to_power_of_two(uint32_t x) {
    x |= x >> 1;
    x |= x >> 2;
    x |= x >> 4;
    x |= x >> 8;
    x |= x >> 16;
    return x ^ (x >> 1);
}

namespace {

// π and √2 as the reference's x87 extended constants (ld.rs:21-23); long double is the x87
// 80-bit format on x86-64 Linux, so each operation rounds once to a 64-bit significand.
const long double PI_LD = 3.141592653589793238462643383279502884L;
const long double SQRT2_LD = 1.414213562373095048801L;

}  // namespace

// nesterov.rs:351 `PI * i / binCntX`, extended precision, to float.
float wavenumber_x(int32_t i, int32_t nx) {
    volatile long double a = PI_LD * (long double)(float)i;
    volatile long double b = a / (long double)(float)nx;
    return (float)b;
}

// nesterov.rs:359 `PI * i / binCntY * binSizeY / binSizeX`, in that order.
float wavenumber_y(int32_t i, int32_t ny, float bin_x, float bin_y) {
    volatile long double a = PI_LD * (long double)(float)i;
    volatile long double b = a / (long double)(float)ny;
    volatile long double c = b * (long double)bin_y;
    volatile long double e = c / (long double)bin_x;
    return (float)e;
}

// nesterov.rs:338 calculateBiVariateNormalCDF (float erf, the √2 a double).
static float bivariate_normal_cdf(float mean_x, float mean_y, float sigma_x, float sigma_y, float lx, float ly, float ux, float uy) {
    double s2 = std::sqrt(2.0);
    float x1 = float(double(mean_x - lx) / (s2 * double(sigma_x)));
    float x2 = float(double(mean_x - ux) / (s2 * double(sigma_x)));
    float y1 = float(double(mean_y - ly) / (s2 * double(sigma_y)));
    float y2 = float(double(mean_y - uy) / (s2 * double(sigma_y)));
    float s = erff(x1) * erff(y1) + erff(x2) * erff(y2) - erff(x1) * erff(y2) - erff(x2) * erff(y1);
    return float(0.25 * double(s));
}

// nesterov.rs:1901 union_area: coordinate compression, exact.
int64_t union_area(const std::vector<Rect>& rects) {
    std::vector<int32_t> xs;
    for (auto& r : rects) {
        xs.push_back(r.lx);
        xs.push_back(r.ux);
    }
    std::sort(xs.begin(), xs.end());
    xs.erase(std::unique(xs.begin(), xs.end()), xs.end());
    int64_t area = 0;
    std::vector<std::pair<int32_t, int32_t>> spans;
    for (size_t w = 0; w + 1 < xs.size(); ++w) {
        int32_t x0 = xs[w], x1 = xs[w + 1];
        spans.clear();
        for (auto& r : rects)
            if (r.lx <= x0 && r.ux >= x1) spans.push_back({r.ly, r.uy});
        std::sort(spans.begin(), spans.end());
        int64_t covered = 0;
        bool have = false;
        int32_t ca = 0, cb = 0;
        for (auto& s : spans) {
            if (have && s.first <= cb)
                cb = std::max(cb, s.second);
            else {
                if (have) covered += int64_t(cb - ca);
                ca = s.first;
                cb = s.second;
                have = true;
            }
        }
        if (have) covered += int64_t(cb - ca);
        area += covered * int64_t(x1 - x0);
    }
    return area;
}

// nesterov.rs:1933 getOverlapArea of a fixed instance and a bin.
static int64_t overlap_area(int32_t blx, int32_t bly, int32_t bux, int32_t buy, const Instance& i, int32_t dbu) {
    int32_t lx = std::max(blx, i.lx), ly = std::max(bly, i.ly), ux = std::min(bux, i.ux), uy = std::min(buy, i.uy);
    if (lx >= ux || ly >= uy) return 0;
    if (i.macro) {
        float f = float(dbu);
        float mean_x = float(i.cx() - i.lx) / f, mean_y = float(i.cy() - i.ly) / f;
        float cdf = bivariate_normal_cdf(mean_x, mean_y, mean_x / 6.0f, mean_y / 6.0f, float(lx - i.lx) / f, float(ly - i.ly) / f,
                                         float(ux - i.lx) / f, float(uy - i.ly) / f);
        float original = float(ux - lx) * float(uy - ly);
        float scaled = cdf * float(i.ux - i.lx) * float(i.uy - i.ly);
        if (scaled >= original) return (int64_t)std::min(scaled, float(double(original) * 1.10));
        return (int64_t)original;
    }
    return int64_t(ux - lx) * int64_t(uy - ly);
}

int32_t init_nesterov(gpl_design* g, std::string* err) {
    const Design& d = g->d;
    Placer& pl = g->pl;
    const gpl_opts& o = g->opts;
    NesterovInit& ni = g->ni;
    ni = NesterovInit{};
    // cells of the common set: one per movable placer instance (no clusters)
    std::vector<int32_t> inst_cell(pl.insts.size(), -1);
    std::vector<Cell> cells;
    for (int32_t k : pl.place) {
        const Instance& i = pl.insts[k];
        inst_cell[k] = int32_t(cells.size());
        Cell c;
        c.inst = k;
        c.macro = i.macro;
        c.locked = i.locked;
        c.lx = c.dlx = i.lx;
        c.ly = c.dly = i.ly;
        c.ux = c.dux = i.ux;
        c.uy = c.duy = i.uy;
        cells.push_back(c);
    }
    // pins as the initial placement left them (before the perturbation below moves the placer's
    // own pins: nesterov.rs:827 copies them first)
    const int32_t np = int32_t(pl.pins.size());
    std::vector<int32_t> pin_xy(size_t(np) * 2), pin_off(size_t(np) * 2), pin_gcell(np);
    for (int32_t p = 0; p < np; ++p) {
        pin_xy[2 * p] = pl.pins[p].cx;
        pin_xy[2 * p + 1] = pl.pins[p].cy;
        pin_off[2 * p] = pl.pins[p].off_x;
        pin_off[2 * p + 1] = pl.pins[p].off_y;
        pin_gcell[p] = pl.pins[p].inst >= 0 ? inst_cell[pl.pins[p].inst] : -1;
    }
    std::vector<std::vector<int32_t>> region_cells;  // per region: common-cell indices
    for (const Base& base : pl.bases) {
        g->log.info(33, strfmt("---- Initialize Nesterov Region: %s", base.group < 0 ? "Top-level" : strfmt("group%d", base.group).c_str()));
        NbInit nb;
        gpl_region& R = nb.r;
        std::memset(&R, 0, sizeof R);
        R.group = base.group;
        R.std_insts_area = base.std_insts_area;
        R.macro_insts_area = base.macro_insts_area;
        R.non_place_insts_area = base.non_place_insts_area;
        // ---- initFillerGCells (nesterov.rs:1690)
        std::vector<int32_t> dxs, dys;
        for (int32_t k : base.place_insts) {
            dxs.push_back(pl.insts[k].dx());
            dys.push_back(pl.insts[k].dy());
        }
        std::sort(dxs.begin(), dxs.end());
        std::sort(dys.begin(), dys.end());
        size_t nn = dxs.size();
        int32_t lo = (int32_t)(double(nn) * 0.05), hi = (int32_t)(double(nn) * 0.95);
        if (lo == hi) {
            lo = 0;
            hi = int32_t(nn);
        }
        if (lo == hi) {
            *err = "a region with no placeable instance (no filler size) is not modelled";
            return GPL_ERR_REFUSED;
        }
        int64_t sx = 0, sy = 0;
        for (int32_t i = lo; i < hi; ++i) {
            sx += dxs[i];
            sy += dys[i];
        }
        R.filler_dx = int32_t(sx / int64_t(hi - lo));
        R.filler_dy = int32_t(sy / int64_t(hi - lo));
        R.white_space_area = base.region_area - base.non_place_insts_area;
        float tmp_target = float(double(float(R.std_insts_area) / float(R.white_space_area - R.macro_insts_area)) + 0.01);
        R.target_density = o.uniform_density ? tmp_target : o.density;
        auto nia = [&]() { return R.std_insts_area + (int64_t)std::round(float(R.macro_insts_area) * R.target_density); };
        int64_t movable = (int64_t)(float(R.white_space_area) * R.target_density);
        int64_t total_filler = movable - nia();
        R.uniform_target_density = float(nia()) / float(R.white_space_area);
        R.uniform_target_density = std::ceil(R.uniform_target_density * 100.0f) / 100.0f;
        if (total_filler < 0) {
            g->log.warn(302, strfmt("Target density %.4f is too low for the available free area.\nAutomatically adjusting to uniform density %.4f.", R.target_density, R.uniform_target_density));
            R.target_density = R.uniform_target_density;
            movable = (int64_t)(float(R.white_space_area) * R.target_density);
            total_filler = movable - nia();
        }
        double scale = std::sqrt(double(total_filler) / (10.0 * double(nia())));
        if (scale > 1.0) {
            int32_t max_x = std::max((int32_t)(double(d.core.dx()) / 1024.0), R.filler_dx);
            int32_t max_y = std::max((int32_t)(double(d.core.dy()) / 1024.0), R.filler_dy);
            R.filler_dx = (int32_t)(double(R.filler_dx) * scale);
            R.filler_dy = (int32_t)(double(R.filler_dy) * scale);
            R.filler_dx = std::min(R.filler_dx, max_x);
            R.filler_dy = std::min(R.filler_dy, max_y);
        }
        R.movable_area = movable;
        R.total_filler_area = total_filler;
        int32_t fa = int32_t(uint32_t(R.filler_dx) * uint32_t(R.filler_dy));
        int32_t count = fa != 0 ? int32_t(total_filler / int64_t(fa)) : 0;
        Mt19937 frng(0);
        Rect r = base.region_bbox;
        std::vector<Cell> fillers;
        for (int32_t i = 0; i < count; ++i) {
            uint32_t rx = frng.next(), ry = frng.next();
            int32_t cx = int32_t(rx % uint32_t(r.dx()) + uint32_t(r.lx));
            int32_t cy = int32_t(ry % uint32_t(r.dy()) + uint32_t(r.ly));
            fillers.push_back(filler_cell(cx, cy, R.filler_dx, R.filler_dy));
        }
        // ---- perturbation (nesterov.rs:1761): per placed instance of the region, in order
        {
            Mt19937 prng(uint32_t(o.seed));
            bool draws = pl.site_y > 0 && o.perturb_dist != 0.0f;
            double rr = o.perturb_dist > 0.0f ? double(o.perturb_dist) * (double(d.dbu) / 1000.0) : std::min(0.5 * double(d.dbu), double(pl.site_y));
            double sigma = rr * 0.3071963263271184;
            for (int32_t k : base.place_insts) {
                int32_t xo = 0, yo = 0;
                if (draws) {
                    double ddx = normal(prng, 0.0, sigma), ddy = normal(prng, 0.0, sigma);
                    xo = (int32_t)std::round(ddx);
                    yo = (int32_t)std::round(ddy);
                }
                Instance& i = pl.insts[k];
                if (!i.locked) {
                    int32_t w = i.dx(), h = i.dy();
                    i.lx += xo;
                    i.ly += yo;
                    i.ux = i.lx + w;
                    i.uy = i.ly + h;
                    int32_t icx = i.cx(), icy = i.cy();
                    for (int32_t p : pl.inst_pins[k]) {
                        pl.pins[p].cx = icx + pl.pins[p].off_x;
                        pl.pins[p].cy = icy + pl.pins[p].off_y;
                    }
                }
                Cell& c = cells[inst_cell[k]];
                c.lx = c.dlx = i.lx;
                c.ly = c.dly = i.ly;
                c.ux = c.dux = i.ux;
                c.uy = c.duy = i.uy;
            }
        }
        std::vector<int32_t> rc;
        for (int32_t k : base.place_insts) rc.push_back(inst_cell[k]);
        R.num_fillers = int32_t(fillers.size());
        for (auto& f : fillers) {
            rc.push_back(int32_t(cells.size()));
            cells.push_back(f);
        }
        // ---- initBins (nesterov.rs:1830)
        int64_t total_bin_area = int64_t(r.ux - r.lx) * int64_t(r.uy - r.ly);
        int64_t avg = base.place_insts.empty() ? 0 : base.place_insts_area / int64_t(base.place_insts.size());
        int64_t ideal_bin_area = R.target_density != 0.0f ? (int64_t)std::round(float(avg) / R.target_density) : 0;
        int32_t ideal_cnt = std::max(ideal_bin_area != 0 ? int32_t(total_bin_area / ideal_bin_area) : 0, 4);
        if (o.bin_cnt > 0) {
            R.cnt_x = R.cnt_y = o.bin_cnt;
        } else {
            int32_t w = r.ux - r.lx, h = r.uy - r.ly;
            int32_t ratio = int32_t(round_down_to_power_of_two(uint32_t(std::max(w, h) / std::min(w, h))));
            int32_t found = 2;
            while (found <= 1024) {
                if ((found == 2 || found * (found * ratio) <= ideal_cnt) && 4 * found * (found * ratio) > ideal_cnt) break;
                found *= 2;
            }
            if (w > h) {
                R.cnt_x = found * ratio;
                R.cnt_y = found;
            } else {
                R.cnt_x = found;
                R.cnt_y = found * ratio;
            }
        }
        g->log.info(28, strfmt("%-21s %7d , %6d", "Bin count (X, Y):", R.cnt_x, R.cnt_y));
        R.lx = r.lx;
        R.ly = r.ly;
        R.ux = r.ux;
        R.uy = r.uy;
        R.size_x = double(r.ux - r.lx) / double(R.cnt_x);
        R.size_y = double(r.uy - r.ly) / double(R.cnt_y);
        nb.bin_x.resize(size_t(R.cnt_x) + 1);
        nb.bin_y.resize(size_t(R.cnt_y) + 1);
        for (int32_t x = 0; x <= R.cnt_x; ++x) nb.bin_x[x] = r.lx + (int32_t)std::round(double(x) * R.size_x);
        for (int32_t y = 0; y <= R.cnt_y; ++y) nb.bin_y[y] = r.ly + (int32_t)std::round(double(y) * R.size_y);
        // ---- updateBinsNonPlaceArea (nesterov.rs:1956)
        {
            std::vector<Instance> fixed;
            for (int32_t k : base.fixed_insts) fixed.push_back(pl.insts[k]);
            for (auto& dm : base.dummies) fixed.push_back(dm);
            const size_t nbins = size_t(R.cnt_x) * size_t(R.cnt_y);
            nb.non_place.assign(nbins, 0);
            nb.non_place_unscaled.assign(nbins, 0);
            auto idx = [&](int32_t lo_, int32_t hi_, int32_t org, double size, int32_t cnt) {
                int32_t a = (int32_t)(double(lo_ - org) / size);
                int32_t b = (int32_t)std::ceil(double(hi_ - org) / size);
                return std::make_pair(std::max(a, 0), std::min(b, cnt));
            };
            std::vector<std::vector<int32_t>> touching(nbins);
            for (size_t k = 0; k < fixed.size(); ++k) {
                const Instance& i = fixed[k];
                if (i.lx >= i.ux || i.ly >= i.uy) continue;
                auto xr = idx(i.lx, i.ux, r.lx, R.size_x, R.cnt_x);
                auto yr = idx(i.ly, i.uy, r.ly, R.size_y, R.cnt_y);
                for (int32_t y = yr.first; y < yr.second; ++y)
                    for (int32_t x = xr.first; x < xr.second; ++x) touching[size_t(y) * R.cnt_x + x].push_back(int32_t(k));
            }
            std::vector<int64_t> uni(nbins, -1), raw(nbins, 0);
            std::vector<Rect> rects;
            for (size_t b = 0; b < nbins; ++b) {
                if (touching[b].empty()) continue;
                int32_t bx = int32_t(b % R.cnt_x), by = int32_t(b / R.cnt_x);
                int32_t blx = nb.bin_x[bx], bux = nb.bin_x[bx + 1], bly = nb.bin_y[by], buy = nb.bin_y[by + 1];
                rects.clear();
                for (int32_t k : touching[b]) {
                    Rect q{std::max(fixed[k].lx, blx), std::max(fixed[k].ly, bly), std::min(fixed[k].ux, bux), std::min(fixed[k].uy, buy)};
                    if (q.lx < q.ux && q.ly < q.uy) rects.push_back(q);
                }
                uni[b] = union_area(rects);
                nb.non_place_unscaled[b] = (int64_t)(float(uni[b]) * R.target_density);
            }
            for (auto& i : fixed) {
                auto xr = idx(i.lx, i.ux, r.lx, R.size_x, R.cnt_x);
                auto yr = idx(i.ly, i.uy, r.ly, R.size_y, R.cnt_y);
                for (int32_t y = yr.first; y < yr.second; ++y)
                    for (int32_t x = xr.first; x < xr.second; ++x)
                        raw[size_t(y) * R.cnt_x + x] += overlap_area(nb.bin_x[x], nb.bin_y[y], nb.bin_x[x + 1], nb.bin_y[y + 1], i, d.dbu);
            }
            for (size_t b = 0; b < nbins; ++b)
                if (raw[b] != 0) nb.non_place[b] += (int64_t)(float(raw[b]) * R.target_density);
            for (size_t b = 0; b < nbins; ++b) {
                if (touching[b].empty()) continue;
                int64_t cap = (int64_t)(float(uni[b]) * R.target_density * 1.10f);
                if (nb.non_place[b] > cap) nb.non_place[b] = cap;
            }
            nb.non_place_raw = raw;
            nb.non_place_union = uni;
        }
        // ---- FFT wavenumbers (nesterov.rs:382)
        nb.wx.resize(R.cnt_x);
        nb.wy.resize(R.cnt_y);
        for (int32_t i = 0; i < R.cnt_x; ++i) nb.wx[i] = wavenumber_x(i, R.cnt_x);
        for (int32_t i = 0; i < R.cnt_y; ++i) nb.wy[i] = wavenumber_y(i, R.cnt_y, float(R.size_x), float(R.size_y));
        // ---- updateDensitySize (nesterov.rs:2022)
        {
            long double thr_x = SQRT2_LD * (long double)R.size_x, thr_y = SQRT2_LD * (long double)R.size_y;
            float size_x = (float)(SQRT2_LD * (long double)(float)R.size_x);
            float size_y = (float)(SQRT2_LD * (long double)(float)R.size_y);
            float thr_xf = (float)thr_x, thr_yf = (float)thr_y;
            for (int32_t gi : rc) {
                Cell& c = cells[gi];
                int32_t dx = c.dx(), dy = c.dy();
                float scale_x, dsx, scale_y, dsy;
                if ((long double)(uint64_t)(int64_t)dx < thr_x) {
                    scale_x = float(dx) / thr_xf;
                    dsx = size_x;
                } else {
                    scale_x = 1.0f;
                    dsx = float(dx);
                }
                if ((long double)(uint64_t)(int64_t)dy < thr_y) {
                    scale_y = float(dy) / thr_yf;
                    dsy = size_y;
                } else {
                    scale_y = 1.0f;
                    dsy = float(dy);
                }
                int32_t ddx = (int32_t)dsx, ddy = (int32_t)dsy;
                int32_t cx = c.d_cx(), cy = c.d_cy();
                c.dlx = cx - ddx / 2;
                c.dly = cy - ddy / 2;
                c.dux = cx + ddx / 2;
                c.duy = cy + ddy / 2;
                c.density_scale = scale_x * scale_y;
            }
        }
        region_cells.push_back(std::move(rc));
        ni.nbs.push_back(std::move(nb));
    }
    // ---- problem arrays in region order
    std::vector<int32_t> new_index(cells.size(), -1);
    int32_t next = 0;
    for (size_t b = 0; b < region_cells.size(); ++b) {
        gpl_region& R = ni.nbs[b].r;
        R.cell_begin = next;
        for (int32_t gi : region_cells[b]) new_index[gi] = next++;
        R.cell_end = next;
    }
    const int32_t ncells = next;
    ni.cell_box.resize(size_t(ncells) * 4);
    ni.cell_dbox.resize(size_t(ncells) * 4);
    ni.density_scale.resize(ncells);
    ni.cell_flags.resize(ncells);
    ni.cell_inst.resize(ncells);
    ni.cell_pin_start.assign(size_t(ncells) + 1, 0);
    std::vector<int32_t> old_of(ncells);
    for (size_t gi = 0; gi < cells.size(); ++gi)
        if (new_index[gi] >= 0) old_of[new_index[gi]] = int32_t(gi);
    for (int32_t i = 0; i < ncells; ++i) {
        const Cell& c = cells[old_of[i]];
        int32_t* bx = &ni.cell_box[size_t(i) * 4];
        int32_t* db = &ni.cell_dbox[size_t(i) * 4];
        bx[0] = c.lx;
        bx[1] = c.ly;
        bx[2] = c.ux;
        bx[3] = c.uy;
        db[0] = c.dlx;
        db[1] = c.dly;
        db[2] = c.dux;
        db[3] = c.duy;
        ni.density_scale[i] = c.density_scale;
        ni.cell_flags[i] = uint8_t((c.inst >= 0 ? GPL_CELL_INSTANCE : 0) | (c.inst >= 0 && c.macro ? GPL_CELL_MACRO : 0) | (c.locked ? GPL_CELL_LOCKED : 0));
        ni.cell_inst[i] = c.inst;
        int32_t np_ = c.inst >= 0 ? int32_t(pl.inst_pins[c.inst].size()) : 0;
        ni.cell_pin_start[i + 1] = ni.cell_pin_start[i] + np_;
    }
    ni.cell_pins.resize(size_t(ni.cell_pin_start[ncells]));
    for (int32_t i = 0; i < ncells; ++i) {
        const Cell& c = cells[old_of[i]];
        if (c.inst < 0) continue;
        int32_t at = ni.cell_pin_start[i];
        for (int32_t p : pl.inst_pins[c.inst]) ni.cell_pins[at++] = p;
    }
    ni.pin_cell.resize(np);
    for (int32_t p = 0; p < np; ++p) ni.pin_cell[p] = pin_gcell[p] >= 0 ? new_index[pin_gcell[p]] : -1;
    ni.pin_offset = pin_off;
    ni.pin_xy = pin_xy;
    ni.pin_net.assign(np, 0);
    ni.net_pin_start.assign(pl.net_pins.size() + 1, 0);
    for (size_t e = 0; e < pl.net_pins.size(); ++e) {
        ni.net_pin_start[e + 1] = ni.net_pin_start[e] + int32_t(pl.net_pins[e].size());
        for (int32_t p : pl.net_pins[e]) {
            ni.net_pins.push_back(p);
            ni.pin_net[p] = int32_t(e);
        }
    }
    ni.net_weight.assign(pl.net_pins.size(), 1.0f);
    ni.regions.clear();
    for (auto& nb : ni.nbs) {
        gpl_region R = nb.r;
        R.non_place = nb.non_place.data();
        R.non_place_unscaled = nb.non_place_unscaled.data();
        R.non_place_raw = nb.non_place_raw.data();
        R.non_place_union = nb.non_place_union.data();
        R.bin_x = nb.bin_x.data();
        R.bin_y = nb.bin_y.data();
        R.wx = nb.wx.data();
        R.wy = nb.wy.data();
        ni.regions.push_back(R);
    }
    ni.inst_box.resize(pl.insts.size() * 4);
    for (size_t k = 0; k < pl.insts.size(); ++k) {
        ni.inst_box[4 * k] = pl.insts[k].lx;
        ni.inst_box[4 * k + 1] = pl.insts[k].ly;
        ni.inst_box[4 * k + 2] = pl.insts[k].ux;
        ni.inst_box[4 * k + 3] = pl.insts[k].uy;
    }
    ni.built = true;
    return GPL_OK;
}

}  // namespace gpl

// ------------------------------------------------------------------------------------------
// C ABI: options and the design
// ------------------------------------------------------------------------------------------
extern "C" {

void gpl_opts_default(gpl_opts* o) {
    std::memset(o, 0, sizeof *o);
    o->density = 0.7f;
    o->init_density_penalty = 0.00008f;
    o->init_wirelength_coef = 0.25f;
    o->min_phi_coef = 0.95f;
    o->max_phi_coef = 1.05f;
    o->overflow = 0.1f;
    o->max_iter = 5000;
    o->seed = 1;
    o->perturb_dist = -1.0f;
    o->initial_place_max_iter = 20;
    o->initial_place_max_fanout = 200;
    o->initial_place_min_diff_length = 1500;
    o->initial_place_max_solver_iter = 100;
    o->initial_place_net_weight_scale = 800.0f;
    // RouteVars::default (route_base.rs:29-45), Opts::default (nesterov.rs:77-78)
    o->routability_end_overflow = 0.3f;
    o->routability_snapshot_overflow = 0.6f;
    o->routability_target_rc = 1.01f;
    o->routability_inflation_ratio_coef = 2.0f;
    o->routability_max_inflation_ratio = 3.0f;
    o->routability_max_density = 0.99f;
    o->routability_min_congestion = 0.95f;
    o->routability_max_inflation_total = 0.30f;
    o->routability_net_weight_max = 1.0f;
    o->routability_congested_nets_percentage = 1.0f;
    o->routability_rc_k[0] = 1.0f;
    o->routability_rc_k[1] = 1.0f;
    o->routability_max_passes = 50;
    o->route_gcells = 64;
    o->route_tracks = 8;
}

int32_t gpl_design_generate(const gpl_gen_config* cfg, const gpl_opts* o, gpl_design** out) {
    if (!cfg || !out) return GPL_ERR_ARG;
    gpl_design* g = new gpl_design();
    g->cfg = *cfg;
    if (o)
        g->opts = *o;
    else
        gpl_opts_default(&g->opts);
    gpl_gen_config& c = g->cfg;
    if (c.degree_p <= 0.0) c.degree_p = 0.5;
    if (c.degree_cap <= 0) c.degree_cap = 64;
    if (c.utilization <= 0.0) c.utilization = 0.7;
    if (c.row_height <= 0) c.row_height = 2000;
    if (c.site_width <= 0) c.site_width = 200;
    if (c.width_min_sites <= 0) c.width_min_sites = 1;
    if (c.width_max_sites <= 0) c.width_max_sites = 8;
    if (c.dbu <= 0) c.dbu = 1000;
    if (c.locality_window <= 0) c.locality_window = 256;
    if (c.macro_rows_min <= 0) c.macro_rows_min = 50;
    if (c.macro_rows_max <= 0) c.macro_rows_max = 400;
    if (c.num_groups > 0) {
        if (!c.group_cells) {
            delete g;
            return GPL_ERR_ARG;
        }
        g->group_cells.assign(c.group_cells, c.group_cells + c.num_groups);
    }
    c.group_cells = g->group_cells.data();
    std::string err;
    int32_t rc = gpl::generate(g, &err);
    if (rc == GPL_OK) rc = gpl::build_placer(g, &err);
    if (rc != GPL_OK) {
        fprintf(stderr, "gpl_design_generate: %s\n", err.c_str());
        delete g;
        return rc;
    }
    *out = g;
    return GPL_OK;
}

void gpl_design_destroy(gpl_design* d) { delete d; }

int32_t gpl_design_initial_place(gpl_design* d) {
    if (!d) return GPL_ERR_ARG;
    std::string err;
    if (d->have_pl_before_init) d->pl = d->pl_before_init;
    d->have_pl_before_init = false;
    int32_t rc = gpl::initial_place(d, &err);
    d->initial_hpwl = gpl::db_hpwl(d);
    return rc;
}

int32_t gpl_design_initial_place_gpu(gpl_design* d) {
    if (!d) return GPL_ERR_ARG;
    d->ip_device = 1;
    const int32_t rc = gpl_design_initial_place(d);
    d->ip_device = 0;
    return rc;
}

int32_t gpl_design_get_info(const gpl_design* d, gpl_design_info* out) {
    if (!d || !out) return GPL_ERR_ARG;
    std::memset(out, 0, sizeof *out);
    out->dbu = d->d.dbu;
    int32_t die[4] = {d->d.die.lx, d->d.die.ly, d->d.die.ux, d->d.die.uy};
    int32_t core[4] = {d->d.core.lx, d->d.core.ly, d->d.core.ux, d->d.core.uy};
    std::memcpy(out->die, die, sizeof die);
    std::memcpy(out->core, core, sizeof core);
    out->num_insts = int32_t(d->pl.insts.size());
    for (auto& i : d->pl.insts) {
        if (i.fixed) {
            out->num_fixed++;
            out->fixed_area += i.area();
        } else {
            out->num_movable++;
            out->movable_area += i.area();
        }
    }
    out->num_nets = int32_t(d->pl.net_pins.size());
    out->num_pins = int32_t(d->pl.pins.size());
    out->num_regions = int32_t(d->pl.bases.size());
    out->initial_hpwl = gpl::db_hpwl(d);
    out->ip_rounds = d->ip_rounds;
    out->ip_error = d->ip_error;
    return GPL_OK;
}

int32_t gpl_init_nesterov_place(gpl_design* d, const gpl_opts* o) {
    if (!d) return GPL_ERR_ARG;
    if (o) d->opts = *o;
    if (d->have_pl_before_init) {
        d->pl = d->pl_before_init;
    } else {
        d->pl_before_init = d->pl;
        d->have_pl_before_init = true;
    }
    std::string err;
    int32_t rc = gpl::init_nesterov(d, &err);
    if (rc != GPL_OK) fprintf(stderr, "gpl_init_nesterov_place: %s\n", err.c_str());
    return rc;
}

void gpl_util_mt19937(uint32_t seed, int32_t n, uint32_t* out) {
    gpl::Mt19937 e(seed);
    for (int32_t i = 0; i < n; ++i) out[i] = e.next();
}

void gpl_util_uniform_int(uint32_t seed, int32_t n, const int32_t* lo, const int32_t* hi, int32_t* out) {
    gpl::Mt19937 e(seed);
    for (int32_t i = 0; i < n; ++i) out[i] = gpl::uniform_int(e, lo[i], hi[i]);
}

uint32_t gpl_util_normal(uint32_t seed, double mean, double sigma, int32_t n, double* out) {
    gpl::Mt19937 e(seed);
    for (int32_t i = 0; i < n; ++i) out[i] = gpl::normal(e, mean, sigma);
    return e.next();
}

float gpl_util_wavenumber_x(int32_t i, int32_t nx) { return gpl::wavenumber_x(i, nx); }
float gpl_util_wavenumber_y(int32_t i, int32_t ny, float bin_x, float bin_y) { return gpl::wavenumber_y(i, ny, bin_x, bin_y); }

int64_t gpl_util_union_area(const int32_t* rects, int32_t n) {
    std::vector<gpl::Rect> v;
    for (int32_t i = 0; i < n; ++i) v.push_back(gpl::Rect{rects[4 * i], rects[4 * i + 1], rects[4 * i + 2], rects[4 * i + 3]});
    return gpl::union_area(v);
}

uint32_t gpl_util_round_down_pow2(uint32_t x) { return gpl::round_down_to_power_of_two(x); }

int32_t gpl_util_struct_sizes(int64_t* out, int32_t cap) {
    const int64_t sz[] = {sizeof(gpl_opts),   sizeof(gpl_gen_config), sizeof(gpl_design_info), sizeof(gpl_region),
                          sizeof(gpl_problem), sizeof(gpl_stats),     sizeof(gpl_iter_row),    sizeof(gpl_kernel_time)};
    const int32_t n = int32_t(sizeof sz / sizeof sz[0]);
    for (int32_t i = 0; i < n && i < cap && out; ++i) out[i] = sz[i];
    return n;
}

int32_t gpl_design_problem(const gpl_design* d, gpl_problem* v) {
    if (!d || !v || !d->ni.built) return GPL_ERR_STATE;
    const gpl::NesterovInit& ni = d->ni;
    std::memset(v, 0, sizeof *v);
    v->die[0] = d->d.die.lx;
    v->die[1] = d->d.die.ly;
    v->die[2] = d->d.die.ux;
    v->die[3] = d->d.die.uy;
    v->num_cells = int32_t(ni.density_scale.size());
    v->cell_box = ni.cell_box.data();
    v->cell_dbox = ni.cell_dbox.data();
    v->density_scale = ni.density_scale.data();
    v->cell_flags = ni.cell_flags.data();
    v->cell_inst = ni.cell_inst.data();
    v->cell_pin_start = ni.cell_pin_start.data();
    v->cell_pins = ni.cell_pins.data();
    v->num_pins = int32_t(ni.pin_cell.size());
    v->pin_cell = ni.pin_cell.data();
    v->pin_offset = ni.pin_offset.data();
    v->pin_xy = ni.pin_xy.data();
    v->pin_net = ni.pin_net.data();
    v->num_nets = int32_t(ni.net_weight.size());
    v->net_pin_start = ni.net_pin_start.data();
    v->net_pins = ni.net_pins.data();
    v->net_weight = ni.net_weight.data();
    v->num_regions = int32_t(ni.regions.size());
    v->regions = ni.regions.data();
    v->opts = d->opts;
    v->dbu = d->d.dbu;
    v->site_x = d->pl.site_x;
    v->site_y = d->pl.site_y;
    v->inst_box = ni.inst_box.data();
    v->num_insts = int32_t(d->pl.insts.size());
    return GPL_OK;
}

}  // extern "C"

// ---- placed locations into the design, and the design as DEF ----------------------------------

extern "C" int32_t gpl_design_write_locations(gpl_design* d, const int32_t* loc, int32_t n, int64_t* hpwl_out) {
    if (!d || (n > 0 && !loc)) return GPL_ERR_ARG;
    gpl::Placer& pl = d->pl;
    gpl::Design& D = d->d;
    for (int32_t i = 0; i < n; ++i) {
        const int32_t k = loc[3 * i], lx = loc[3 * i + 1], ly = loc[3 * i + 2];
        if (k < 0 || k >= (int32_t)pl.insts.size()) return GPL_ERR_ARG;
        const int32_t di = pl.insts[k].db;
        if (di < 0) continue;  // a dummy: no design instance
        gpl::Rect& b = D.insts[di].bbox;
        const int32_t w = b.dx(), h = b.dy();
        b = gpl::Rect{lx, ly, lx + w, ly + h};
        D.insts[di].placed = true;
        if (di < (int32_t)pl.db_bbox.size()) pl.db_bbox[di] = b;
    }
    if (hpwl_out) {
        // the nets' terminal boxes (db_router.rs:98-112): each instance terminal at its instance's
        // centre plus its offset, each port its box
        int64_t total = 0;
        for (auto& net : D.nets) {
            int32_t lx = INT32_MAX, ly = INT32_MAX, ux = INT32_MIN, uy = INT32_MIN;
            for (auto& t : net.iterms) {
                const gpl::Rect& b = D.insts[t.inst].bbox;
                const int32_t x = (b.lx + b.ux) / 2 + t.off_x, y = (b.ly + b.uy) / 2 + t.off_y;
                lx = std::min(lx, x);
                ly = std::min(ly, y);
                ux = std::max(ux, x);
                uy = std::max(uy, y);
            }
            for (auto& q : net.bterms) {
                lx = std::min(lx, q.bbox.lx);
                ly = std::min(ly, q.bbox.ly);
                ux = std::max(ux, q.bbox.ux);
                uy = std::max(uy, q.bbox.uy);
            }
            if (ux >= lx && uy >= ly) total += int64_t(ux - lx) + int64_t(uy - ly);
        }
        *hpwl_out = total;
    }
    return GPL_OK;
}

extern "C" int32_t gpl_design_bases(const gpl_design* d, int32_t* out, int32_t cap) {
    if (!d) return GPL_ERR_ARG;
    const int32_t n = (int32_t)d->pl.bases.size();
    for (int32_t b = 0; b < n && out && b < cap; ++b) {
        const gpl::Base& B = d->pl.bases[b];
        out[4 * b] = B.group;
        out[4 * b + 1] = (int32_t)B.place_insts.size();
        out[4 * b + 2] = (int32_t)B.fixed_insts.size();
        out[4 * b + 3] = (int32_t)B.dummies.size();
    }
    return n;
}

// DEF 5.8 of the design as it stands: the die, the rows, every instance (PLACED where the placer
// moved it, FIXED where it is fixed), the IO ports with their boxes and every net's terminals.  The
// synthetic designs have no library: a master is named by its size (W<sites>H<rows> for standard
// cells, BLOCK_<w>x<h> for blocks) and an instance terminal by its place among the instance's.
extern "C" int32_t gpl_design_write_def(const gpl_design* d, const char* path) {
    if (!d || !path) return GPL_ERR_ARG;
    FILE* f = std::fopen(path, "w");
    if (!f) return GPL_ERR_ARG;
    const gpl::Design& D = d->d;
    const int32_t sx = std::max(1, D.site_x), sy = std::max(1, D.site_y);
    std::fprintf(f, "VERSION 5.8 ;\nDIVIDERCHAR \"/\" ;\nBUSBITCHARS \"[]\" ;\nDESIGN synthetic_%u ;\n", d->cfg.seed);
    std::fprintf(f, "UNITS DISTANCE MICRONS %d ;\n\n", D.dbu);
    std::fprintf(f, "DIEAREA ( %d %d ) ( %d %d ) ;\n\n", D.die.lx, D.die.ly, D.die.ux, D.die.uy);
    for (size_t r = 0; r < D.rows.size(); ++r) {
        const gpl::Rect& R = D.rows[r];
        std::fprintf(f, "ROW ROW_%zu core_site %d %d %s DO %d BY 1 STEP %d 0 ;\n", r, R.lx, R.ly, (r % 2) ? "FS" : "N", R.dx() / sx, sx);
    }
    std::fprintf(f, "\nCOMPONENTS %zu ;\n", D.insts.size());
    for (size_t i = 0; i < D.insts.size(); ++i) {
        const gpl::DbInst& I = D.insts[i];
        const gpl::Rect& b = i < d->pl.db_bbox.size() ? d->pl.db_bbox[i] : I.bbox;
        char master[64];
        if (I.block)
            std::snprintf(master, sizeof master, "BLOCK_%dx%d", b.dx(), b.dy());
        else
            std::snprintf(master, sizeof master, "W%dH%d", b.dx() / sx, std::max(1, b.dy() / sy));
        std::fprintf(f, "- i%zu %s + %s ( %d %d ) N ;\n", i, master, I.fixed ? "FIXED" : "PLACED", b.lx, b.ly);
    }
    std::fprintf(f, "END COMPONENTS\n\n");
    size_t nports = 0;
    for (auto& n : D.nets) nports += n.bterms.size();
    std::fprintf(f, "PINS %zu ;\n", nports);
    size_t port = 0;
    for (size_t e = 0; e < D.nets.size(); ++e)
        for (auto& p : D.nets[e].bterms) {
            const int32_t cx = (p.bbox.lx + p.bbox.ux) / 2, cy = (p.bbox.ly + p.bbox.uy) / 2;
            std::fprintf(f, "- io%zu + NET n%zu + DIRECTION INOUT + USE SIGNAL\n  + LAYER metal1 ( %d %d ) ( %d %d )\n  + FIXED ( %d %d ) N ;\n",
                         port++, e, p.bbox.lx - cx, p.bbox.ly - cy, p.bbox.ux - cx, p.bbox.uy - cy, cx, cy);
        }
    std::fprintf(f, "END PINS\n\nNETS %zu ;\n", D.nets.size());
    port = 0;
    for (size_t e = 0; e < D.nets.size(); ++e) {
        std::fprintf(f, "- n%zu", e);
        for (auto& t : D.nets[e].iterms) std::fprintf(f, " ( i%d p%d )", t.inst, t.order);
        for (size_t q = 0; q < D.nets[e].bterms.size(); ++q) std::fprintf(f, " ( PIN io%zu )", port++);
        std::fprintf(f, " ;\n");
    }
    std::fprintf(f, "END NETS\n\nEND DESIGN\n");
    const bool ok = std::ferror(f) == 0;
    return std::fclose(f) == 0 && ok ? GPL_OK : GPL_ERR_ARG;
}
