// Host-side model of the placer: design records, the placer's instances/pins/nets, initial
// placement and the Nesterov placer's construction.  Everything here runs once per placement on
// the host; the per-iteration work is in nesterov.cpp / kernels.hip.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/gpl_hip.h"

namespace gpl {

// ------------------------------------------------------------------------------------------
// Log: the reference's coded line format (`[INFO GPL-0001] text`, log.rs:53-107) and report
// text.  Lines are kept in memory; a 1000-per-code print limit as the reference's logger has.
// ------------------------------------------------------------------------------------------
class Log {
public:
    void info(int id, const std::string& msg) { coded("INFO", id, msg); }
    void warn(int id, const std::string& msg) { coded("WARNING", id, msg); }
    void error(int id, const std::string& msg) { coded("ERROR", id, msg); }
    void report(const std::string& msg) { lines_.push_back(msg); }
    const std::vector<std::string>& lines() const { return lines_; }
    std::string joined() const;
    bool echo = false;  // also print to stderr as written

private:
    void coded(const char* level, int id, const std::string& msg);
    std::vector<std::string> lines_;
    std::vector<std::pair<int, int>> counts_;
};

std::string strfmt(const char* fmt, ...) __attribute__((format(printf, 1, 2)));

// ------------------------------------------------------------------------------------------
// RNG: the 32-bit Mersenne Twister (normal.rs:38-120) and boost's uniform_int (normal.rs:127).
// ------------------------------------------------------------------------------------------
class Mt19937 {
public:
    explicit Mt19937(uint32_t seed);
    uint32_t next();

private:
    void twist();
    uint32_t x_[624];
    int i_;
};

int32_t uniform_int(Mt19937& eng, int32_t lo, int32_t hi);
// One draw of boost's normal_distribution<double>(mean, sigma) from the engine: its ziggurat unit
// normal times sigma plus mean (normal.rs:120-124), word for word as the reference draws it.
double normal(Mt19937& eng, double mean, double sigma);
double uniform01(Mt19937& eng);  // 53-bit uniform in [0, 1)

// ------------------------------------------------------------------------------------------
// Design records (placer_base.rs:18-206), reduced to what the synthetic generator produces.
// ------------------------------------------------------------------------------------------
struct Rect {
    int32_t lx = 0, ly = 0, ux = 0, uy = 0;
    int32_t dx() const { return ux - lx; }
    int32_t dy() const { return uy - ly; }
    int64_t area() const { return int64_t(dx()) * int64_t(dy()); }
};

struct DbInst {
    Rect bbox;          // lower-left is the location
    bool fixed = false;
    bool block = false; // master is a BLOCK
    int32_t group = -1;
    int32_t signal_pins = 0;
    bool placed = false;
};

struct Term {
    int32_t inst;
    int32_t off_x, off_y;  // the pin shape's centre less the master centre (R0)
    int32_t order;         // position among the instance's terminals
};

struct Port {
    Rect bbox;
};

struct DbNet {
    std::vector<Term> iterms;
    std::vector<Port> bterms;
};

struct Region {
    Rect box;
};

struct Design {
    int32_t dbu = 1000;
    Rect die, core;
    int32_t site_x = 200, site_y = 2000;
    std::vector<Rect> rows;
    std::vector<DbInst> insts;
    std::vector<DbNet> nets;
    std::vector<Region> regions;  // group g has region g
};

// ------------------------------------------------------------------------------------------
// The placer (placer_base.rs:219-300, 420-446; initial_place.rs:36-70)
// ------------------------------------------------------------------------------------------
struct Instance {
    int32_t db = -1;   // design instance, -1 for a dummy
    int32_t lx = 0, ly = 0, ux = 0, uy = 0;
    bool fixed = false;
    bool macro = false;
    bool locked = false;
    int32_t dx() const { return ux - lx; }
    int32_t dy() const { return uy - ly; }
    int32_t cx() const { return (lx + ux) / 2; }
    int32_t cy() const { return (ly + uy) / 2; }
    int64_t area() const { return int64_t(dx()) * int64_t(dy()); }
};

struct Base {
    int32_t group = -1;
    Rect region_bbox;
    int64_t region_area = 0;
    std::vector<int32_t> place_insts;  // into Placer::insts
    std::vector<int32_t> fixed_insts;
    std::vector<Instance> dummies;
    int64_t place_insts_area = 0, non_place_insts_area = 0, std_insts_area = 0, macro_insts_area = 0;
};

struct Pin {
    int32_t inst = -1;  // placer instance, -1 for a port
    bool port = false;
    int32_t off_x = 0, off_y = 0, cx = 0, cy = 0;
    bool min_x = false, max_x = false, min_y = false, max_y = false;
};

struct Placer {
    int32_t site_x = 0, site_y = 0;
    std::vector<Instance> insts;
    std::vector<Base> bases;
    std::vector<Pin> pins;
    std::vector<std::vector<int32_t>> net_pins;
    std::vector<std::vector<int32_t>> inst_pins;
    std::vector<int32_t> place;  // movable placer instances
    std::vector<Rect> db_bbox;   // each design instance's database box
};

// ------------------------------------------------------------------------------------------
// The Nesterov placer's state as init_nesterov_place leaves it (nesterov.rs:787-933)
// ------------------------------------------------------------------------------------------
struct NbInit {
    gpl_region r{};
    std::vector<int64_t> non_place, non_place_unscaled;
    std::vector<int64_t> non_place_raw, non_place_union;  // the fixed instances' overlap sums and union areas (-1: none touches)
    std::vector<int32_t> bin_x, bin_y;
    std::vector<float> wx, wy;
};

struct NesterovInit {
    // problem cells in region order: each region's instances, then its fillers
    std::vector<int32_t> cell_box, cell_dbox, cell_inst, cell_pin_start, cell_pins;
    std::vector<float> density_scale;
    std::vector<uint8_t> cell_flags;
    std::vector<int32_t> pin_cell, pin_offset, pin_xy, pin_net;
    std::vector<int32_t> net_pin_start, net_pins;
    std::vector<float> net_weight;
    std::vector<NbInit> nbs;
    std::vector<gpl_region> regions;
    std::vector<int32_t> inst_box;
    bool built = false;
};

}  // namespace gpl

struct gpl_design {
    gpl_gen_config cfg{};
    gpl_opts opts{};
    gpl::Design d;
    gpl::Placer pl;
    // the placer as it stood before the first init_nesterov (which perturbs it): restored by every
    // later one, so that re-initialising with another seed starts from the same placement
    gpl::Placer pl_before_init;
    bool have_pl_before_init = false;
    gpl::NesterovInit ni;
    gpl::Log log;
    int32_t ip_rounds = 0;
    float ip_error = 0.0f;
    int64_t initial_hpwl = 0;
    std::vector<int32_t> group_cells;
    int32_t ip_device = 0;  // 1: the initial placement's rounds on the GPU (gpl_design_initial_place_gpu)
};

namespace gpl {
// One BiCGSTAB solve's outcome (bicgstab.rs:143-197).
struct IpSolve {
    int32_t iterations;
    float error;
    bool converged;
};
// One initial-placement round's pin flags, B2B system and both solves on the device, bit for bit
// as the host's (ip_gpu.hip); x, y in and out.
int32_t ip_round_device(gpl_design* g, const std::vector<int32_t>& ext, std::vector<float>& x, std::vector<float>& y, IpSolve* ex,
                        IpSolve* ey, std::string* err);
// placer_base.rs:301 placer_base_common + 430 placer_base + initial_place.rs:100 pins
int32_t build_placer(gpl_design* g, std::string* err);
// initial_place.rs:420
int32_t initial_place(gpl_design* g, std::string* err);
// nesterov.rs:787
int32_t init_nesterov(gpl_design* g, std::string* err);
// The database-box HPWL (initial_place.rs:381)
int64_t db_hpwl(const gpl_design* g);
float wavenumber_x(int32_t i, int32_t nx);
float wavenumber_y(int32_t i, int32_t ny, float bin_x, float bin_y);
int64_t union_area(const std::vector<Rect>& rects);
uint32_t round_down_to_power_of_two(uint32_t x);
}  // namespace gpl
