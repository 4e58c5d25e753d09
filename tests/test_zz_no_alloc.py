"""gpl_density_scatter_noalloc (the density map with no allocation, its scratch in the library's
fixed device array) against gpl_density_scatter (the context's entry, same kernels) through the C
ABI: two calls in a row, the second on a smaller problem (fewer cells, a smaller grid), each
bit-exact in the integer areas and the density and exact in the FP64 overflows; output buffers
start as garbage (the entry clears them).  The device buffers come from the test's own hipMalloc
calls (the test allocates; the entry point does not).  Refusals (a grid above 4096 bins a side, a
null argument) are checked on the host before any launch."""
import ctypes as C

import numpy as np
import pytest

import gpl_amd as G
from gpl_amd import _ffi as F
from tests.test_gpu_parity import _ctx
from tests.tinyproblem import Tiny

pytestmark = pytest.mark.gpu


class _Hip:
    """The HIP runtime the library itself links (torch bundles another one, which cannot share a
    process with it), for the test's own device buffers."""

    def __init__(self):
        F.lib()
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(a.nbytes, 1)) == 0
        self.bufs.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p.value

    def get(self, ptr, n, dtype):
        a = np.empty(n, dtype)
        assert self.L.hipDeviceSynchronize() == 0
        assert self.L.hipMemcpy(a.ctypes.data, C.c_void_p(ptr), a.nbytes, 2) == 0  # device to host
        return a

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _noalloc(dbox, scale, flags, R, bin_x, bin_y, non_place, non_place_unscaled):
    M = R["cnt_x"] * R["cnt_y"]
    h = _Hip()
    try:
        io = F.DensityIO()
        io.num_cells = len(scale)
        io.dbox = h.put(np.asarray(dbox, np.int32).reshape(-1))
        io.density_scale = h.put(np.asarray(scale, np.float32))
        io.cell_flags = h.put(np.asarray(flags, np.uint8))
        io.lx, io.ly, io.ux, io.uy = R["lx"], R["ly"], R["ux"], R["uy"]
        io.cnt_x, io.cnt_y, io.size_x, io.size_y = R["cnt_x"], R["cnt_y"], R["size_x"], R["size_y"]
        io.target_density = R["target_density"]
        io.bin_x = h.put(np.asarray(bin_x, np.int32))
        io.bin_y = h.put(np.asarray(bin_y, np.int32))
        io.non_place_f = h.put(np.stack([np.asarray(non_place, np.int64).astype(np.float32),
                                         np.asarray(non_place_unscaled, np.int64).astype(np.float32)], 1).reshape(-1))
        # outputs start as garbage: the entry clears what it accumulates into
        io.inst_area = h.put(np.full(M, 0x5A5A5A5A, np.int64))
        io.filler_area = h.put(np.full(M, -77, np.int64))
        io.density = h.put(np.full(M, np.nan, np.float32))
        io.overflow = h.put(np.full(2, -1.0, np.float64))
        io.stream = None
        rc = F.lib().gpl_density_scatter_noalloc(C.byref(io))
        out = (rc, h.get(io.inst_area, M, np.int64), h.get(io.filler_area, M, np.int64), h.get(io.density, M, np.float32),
               h.get(io.overflow, 2, np.float64))
    finally:
        h.free()
    return out


def _region(a, b):
    R = dict(a["regions"][b])
    lo, hi = R["cell_begin"], R["cell_end"]
    db = a["cell_dbox"][lo:hi]
    return R, db, a["density_scale"][lo:hi], a["cell_flags"][lo:hi]


def _check(g, R, db, scale, flags):
    assert np.all((db[:, 2] - db[:, 0]) % 2 == 0) and np.all((db[:, 3] - db[:, 1]) % 2 == 0)
    xy = np.stack([(db[:, 0] + db[:, 2]) // 2, (db[:, 1] + db[:, 3]) // 2], 1).astype(np.float32)
    g.set_coordinates(0, G.GPL_SLOT_CUR_SLP, xy)
    inst, fill, dens, ovf = g.density_scatter(0, G.GPL_SLOT_CUR_SLP)
    rc, ni, nf, nd, no = _noalloc(db, scale, flags, R, R["bin_x"], R["bin_y"], R["non_place"], R["non_place_unscaled"])
    assert rc == 0
    assert np.array_equal(ni, inst) and np.array_equal(nf, fill)
    assert np.array_equal(nd, dens)
    assert no[0] == ovf[0] and no[1] == ovf[1]
    assert inst.sum() > 0


def test_noalloc_density_map_matches_the_context_entry_twice():
    # first call: the 2,000-cell workload (64 x 64 bins)
    d = G.Design.from_config("c1")
    d.initial_place()
    d.init_nesterov_place()
    a = d.arrays()
    R, db, sc, fl = _region(a, 0)
    g = G.NesterovGPU(d)
    _check(g, R, db, sc, fl)
    g.close()
    d.close()
    # second call, smaller: 300 cells (fillers and macros among them) on 16 x 16 bins
    rng = np.random.default_rng(2)
    cells = []
    for i in range(300):
        # even sizes: centring a box at its own centre (set_density_center, nesterov.rs:2139) then
        # leaves it as it is, so both entries scatter the same boxes
        w, h = 2 * ((rng.uniform(0.3, 2.0, 2) * 250).astype(int) + 1)
        x, y = rng.uniform(500, 7500, 2).astype(int)
        cells.append(dict(box=(int(x), int(y), int(x + w), int(y + h)), scale=float(rng.uniform(0.4, 1.0)),
                          flags=0 if i % 9 == 0 else (1 | (2 if i % 31 == 0 else 0))))
    t = Tiny(cells, [], regions=[dict(lx=0, ly=0, ux=8000, uy=8000, cnt_x=16, cnt_y=16, target_density=0.75,
                                      non_place=rng.integers(0, 50000, 256))])
    a = dict(cell_dbox=t.a["cell_dbox"].reshape(-1, 4), density_scale=t.a["density_scale"], cell_flags=t.a["cell_flags"])
    p = t.problem.regions[0]
    M = 256
    R = dict(lx=p.lx, ly=p.ly, ux=p.ux, uy=p.uy, cnt_x=p.cnt_x, cnt_y=p.cnt_y, size_x=p.size_x, size_y=p.size_y,
             target_density=p.target_density, cell_begin=0, cell_end=300,
             bin_x=np.ctypeslib.as_array(p.bin_x, (17,)).copy(), bin_y=np.ctypeslib.as_array(p.bin_y, (17,)).copy(),
             non_place=np.ctypeslib.as_array(p.non_place, (M,)).copy(),
             non_place_unscaled=np.ctypeslib.as_array(p.non_place_unscaled, (M,)).copy())
    g = _ctx(t)
    _check(g, R, a["cell_dbox"], a["density_scale"], a["cell_flags"])
    g.close()


def test_noalloc_refuses_what_it_cannot_hold():
    R = dict(lx=0, ly=0, ux=5000, uy=5000, cnt_x=5000, cnt_y=1, size_x=1.0, size_y=5000.0, target_density=1.0)
    io = F.DensityIO()
    io.cnt_x, io.cnt_y, io.size_x, io.size_y = R["cnt_x"], R["cnt_y"], R["size_x"], R["size_y"]
    assert F.lib().gpl_density_scatter_noalloc(C.byref(io)) == F.GPL_ERR_ARG  # 5000 bins a side
    assert F.lib().gpl_density_scatter_noalloc(None) == F.GPL_ERR_ARG
