"""Hand-built Nesterov problems for known-answer tests: a gpl_problem (ctypes) over numpy arrays
the returned object keeps alive.  One region unless several are given."""
import ctypes as C

import numpy as np

import gpl_amd as G
from gpl_amd import _ffi as F


class Tiny:
    def __init__(self, cells, nets, ports=(), regions=None, opts=None):
        """cells: dicts box=(lx,ly,ux,uy), dbox (default box), scale (1.0), flags (instance),
        pins=[(off_x, off_y), ...]; nets: lists of pin ids (cell pins numbered in cell order, then
        the ports); ports: fixed pin positions (x, y); regions: dicts lx,ly,ux,uy,cnt_x,cnt_y,
        target_density, cells=(begin, end), non_place (optional, per bin)."""
        N = len(cells)
        pin_cell, pin_off, pin_xy = [], [], []
        cell_pin_start = [0]
        for g, c in enumerate(cells):
            for (ox, oy) in c.get("pins", []):
                pin_cell.append(g)
                pin_off.append((ox, oy))
                b = c["box"]
                pin_xy.append(((b[0] + b[2]) // 2 + ox, (b[1] + b[3]) // 2 + oy))
            cell_pin_start.append(len(pin_cell))
        for (x, y) in ports:
            pin_cell.append(-1)
            pin_off.append((0, 0))
            pin_xy.append((x, y))
        Pn = len(pin_cell)
        pin_net = [0] * Pn
        net_start, net_pins = [0], []
        for e, ps in enumerate(nets):
            for q in ps:
                pin_net[q] = e
                net_pins.append(q)
            net_start.append(len(net_pins))
        self.a = dict(
            cell_box=np.array([c["box"] for c in cells], np.int32).reshape(-1),
            cell_dbox=np.array([c.get("dbox", c["box"]) for c in cells], np.int32).reshape(-1),
            density_scale=np.array([c.get("scale", 1.0) for c in cells], np.float32),
            cell_flags=np.array([c.get("flags", G.GPL_CELL_INSTANCE) for c in cells], np.uint8),
            cell_inst=np.arange(N, dtype=np.int32),
            cell_pin_start=np.array(cell_pin_start, np.int32),
            cell_pins=np.arange(cell_pin_start[-1], dtype=np.int32),
            pin_cell=np.array(pin_cell, np.int32), pin_offset=np.array(pin_off, np.int32).reshape(-1),
            pin_xy=np.array(pin_xy, np.int32).reshape(-1), pin_net=np.array(pin_net, np.int32),
            net_pin_start=np.array(net_start, np.int32), net_pins=np.array(net_pins, np.int32),
            net_weight=np.ones(len(nets), np.float32),
            inst_box=np.array([c["box"] for c in cells], np.int32).reshape(-1),
        )
        if regions is None:
            raise ValueError("give the region")
        L = F.lib()
        self.regs = (F.Region * len(regions))()
        self._keep = []
        for b, r in enumerate(regions):
            R = self.regs[b]
            R.lx, R.ly, R.ux, R.uy = r["lx"], r["ly"], r["ux"], r["uy"]
            R.cnt_x, R.cnt_y = r["cnt_x"], r["cnt_y"]
            R.size_x = (r["ux"] - r["lx"]) / r["cnt_x"]
            R.size_y = (r["uy"] - r["ly"]) / r["cnt_y"]
            R.target_density = r.get("target_density", 1.0)
            R.uniform_target_density = R.target_density
            R.cell_begin, R.cell_end = r.get("cells", (0, N))
            R.num_fillers = r.get("num_fillers", 0)
            M = R.cnt_x * R.cnt_y
            npl = np.ascontiguousarray(r.get("non_place", np.zeros(M)), np.int64)
            npu = np.ascontiguousarray(r.get("non_place_unscaled", npl), np.int64)
            bx = (R.lx + np.floor(np.arange(R.cnt_x + 1) * R.size_x + 0.5)).astype(np.int32)
            by = (R.ly + np.floor(np.arange(R.cnt_y + 1) * R.size_y + 0.5)).astype(np.int32)
            wx = np.array([L.gpl_util_wavenumber_x(i, R.cnt_x) for i in range(R.cnt_x)], np.float32)
            wy = np.array([L.gpl_util_wavenumber_y(i, R.cnt_y, np.float32(R.size_x), np.float32(R.size_y)) for i in range(R.cnt_y)], np.float32)
            self._keep += [npl, npu, bx, by, wx, wy]
            R.non_place = npl.ctypes.data_as(C.POINTER(C.c_int64))
            R.non_place_unscaled = npu.ctypes.data_as(C.POINTER(C.c_int64))
            R.bin_x = bx.ctypes.data_as(C.POINTER(C.c_int32))
            R.bin_y = by.ctypes.data_as(C.POINTER(C.c_int32))
            R.wx = wx.ctypes.data_as(C.POINTER(C.c_float))
            R.wy = wy.ctypes.data_as(C.POINTER(C.c_float))
            R.std_insts_area = int(sum((c["box"][2] - c["box"][0]) * (c["box"][3] - c["box"][1]) for c in cells[R.cell_begin:R.cell_end]
                                       if c.get("flags", 1) & 1))
            R.white_space_area = (R.ux - R.lx) * (R.uy - R.ly)
            R.group = -1 if b == 0 else b
        p = F.Problem()
        a = self.a
        p.num_cells, p.num_pins, p.num_nets = N, Pn, len(nets)
        i32p, f32p = C.POINTER(C.c_int32), C.POINTER(C.c_float)
        for k in ("cell_box", "cell_dbox", "cell_inst", "cell_pin_start", "cell_pins", "pin_cell", "pin_offset", "pin_xy",
                  "pin_net", "net_pin_start", "net_pins", "inst_box"):
            setattr(p, k, a[k].ctypes.data_as(i32p))
        p.density_scale = a["density_scale"].ctypes.data_as(f32p)
        p.net_weight = a["net_weight"].ctypes.data_as(f32p)
        p.cell_flags = a["cell_flags"].ctypes.data_as(C.POINTER(C.c_uint8))
        p.num_regions = len(regions)
        p.regions = C.cast(self.regs, C.POINTER(F.Region))
        p.opts = opts if opts is not None else G.default_opts()
        p.dbu, p.site_x, p.site_y = 1000, 200, 2000
        p.num_insts = N
        self.problem = p


def two_by_two(cells, nets=(), ports=(), target_density=1.0, non_place=None):
    """A 200 x 200 region of 2 x 2 bins of 100 x 100."""
    reg = dict(lx=0, ly=0, ux=200, uy=200, cnt_x=2, cnt_y=2, target_density=target_density)
    if non_place is not None:
        reg["non_place"] = non_place
    return Tiny(cells, list(nets), ports, [reg])
