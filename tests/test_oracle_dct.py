"""The oracle's transforms and Poisson solve, pinned by brute-force sums, scipy.fft, the
forward/inverse identity and the analytic single-mode solution of the Poisson problem."""
import numpy as np
import pytest
import scipy.fft as sf

import oracle as O


def _run(fn, x, dtype):
    a = np.array(x, dtype=dtype, copy=True)
    getattr(O.lib(), fn)(len(a), O.f64(a) if dtype == np.float64 else O.f32(a))
    return a


@pytest.mark.parametrize("n", [2, 4, 8, 16, 64, 256, 6, 12, 24, 48, 96, 384, 768, 10])
def test_1d_transforms_match_brute_force(n):
    """Powers of two, 3 * 2^k (the radix-3 split of the half-length FFT) and a length neither
    covers (10: the direct sums themselves)."""
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    for fast, direct in [("orc_dct2_f64", "orc_direct_dct2_f64"), ("orc_dct3_f64", "orc_direct_dct3_f64"),
                         ("orc_dst3_f64", "orc_direct_dst3_f64")]:
        a, b = _run(fast, x, np.float64), _run(direct, x, np.float64)
        assert np.max(np.abs(a - b)) < 1e-11 * n, (fast, n)
    for f32, f64 in [("orc_dct2_f32", "orc_dct2_f64"), ("orc_dct3_f32", "orc_dct3_f64"), ("orc_dst3_f32", "orc_dst3_f64")]:
        a, b = _run(f32, x, np.float32), _run(f64, x, np.float64)
        assert np.max(np.abs(a - b)) < 2e-6 * n * np.max(np.abs(x)), (f32, n)


@pytest.mark.parametrize("n", [4, 32, 128, 12, 96, 768])
def test_1d_transforms_match_scipy(n):
    rng = np.random.default_rng(100 + n)
    x = rng.standard_normal(n)
    # Ooura's DCT-II is scipy's unnormalised type 2 over 2
    assert np.allclose(_run("orc_dct2_f64", x, np.float64), sf.dct(x, type=2) / 2, atol=1e-12)
    # Ooura's DCT-III (a[0] at full weight) is scipy's type 3 of (2 a[0], a[1:]) over 2
    x2 = x.copy()
    x2[0] *= 2
    assert np.allclose(_run("orc_dct3_f64", x, np.float64), sf.dct(x2, type=3) / 2, atol=1e-12)
    # Ooura's DST-III with A[n] in a[0] is scipy's type 3 of (A[1..n-1], 2 A[n]) over 2
    xs = np.concatenate([x[1:], [2 * x[0]]])
    assert np.allclose(_run("orc_dst3_f64", x, np.float64), sf.dst(xs, type=3) / 2, atol=1e-12)


def test_forward_inverse_identity():
    n = 64
    x = np.random.default_rng(7).standard_normal(n)
    c = _run("orc_dct2_f64", x, np.float64)
    c[0] *= 0.5
    y = _run("orc_dct3_f64", c, np.float64) * 2.0 / n
    assert np.allclose(y, x, atol=1e-12)


def _wavenumbers(nx, ny, bx, by):
    import gpl_amd._ffi as F
    L = F.lib()
    wx = np.array([L.gpl_util_wavenumber_x(i, nx) for i in range(nx)], np.float32)
    wy = np.array([L.gpl_util_wavenumber_y(i, ny, bx, by) for i in range(ny)], np.float32)
    return wx, wy


@pytest.mark.parametrize("nx,ny,bx,by", [(16, 16, 3.0, 3.0), (32, 16, 2.0, 3.0), (8, 32, 5.0, 1.25), (24, 48, 2.0, 1.0), (96, 96, 1.0, 1.0)])
@pytest.mark.parametrize("u,v", [(1, 1), (3, 2), (0, 5), (4, 0), (7, 7)])
def test_poisson_single_mode_is_an_eigenfunction(nx, ny, bx, by, u, v):
    """rho = cos(pi u (x+1/2)/nx) cos(pi v (y+1/2)/ny) must give phi = rho / (wx_u^2 + wy_v^2),
    field_x = wx_u / (wx_u^2 + wy_v^2) sin(..x..) cos(..y..), field_y likewise with the roles
    swapped -- pinning the halving, the 4/(nx ny) scale, the divide and the inverse conventions."""
    if u >= nx or v >= ny:
        pytest.skip("mode outside the grid")
    wx, wy = _wavenumbers(nx, ny, bx, by)
    X, Y = np.meshgrid(np.arange(nx), np.arange(ny))  # y-major: [y][x]
    cx, sx = np.cos(np.pi * u * (X + 0.5) / nx), np.sin(np.pi * u * (X + 0.5) / nx)
    cy, sy = np.cos(np.pi * v * (Y + 0.5) / ny), np.sin(np.pi * v * (Y + 0.5) / ny)
    rho = (cx * cy).astype(np.float32).reshape(-1)
    w2 = float(wx[u]) ** 2 + float(wy[v]) ** 2
    phi, fx, fy = O.solve_f64(nx, ny, wx, wy, rho)
    r64 = (cx * cy).reshape(-1)
    assert np.allclose(phi, r64 / w2, atol=1e-9 / w2 + 1e-6 * np.max(np.abs(r64 / w2)))
    assert np.allclose(fx, (float(wx[u]) / w2 * sx * cy).reshape(-1), atol=1e-6 / np.sqrt(w2))
    assert np.allclose(fy, (float(wy[v]) / w2 * cx * sy).reshape(-1), atol=1e-6 / np.sqrt(w2))
    # the float path agrees with the double one to float rounding
    p32, fx32, fy32 = O.solve_f32(nx, ny, wx, wy, rho)
    scale = max(np.max(np.abs(phi)), 1e-30)
    assert np.max(np.abs(p32 - phi)) < 1e-4 * scale


def test_solve_is_linear_and_zero_mean_blind():
    nx, ny = 32, 32
    wx, wy = _wavenumbers(nx, ny, 1.0, 1.0)
    rng = np.random.default_rng(3)
    a, b = rng.random(nx * ny).astype(np.float32), rng.random(nx * ny).astype(np.float32)
    pa = O.solve_f64(nx, ny, wx, wy, a)
    pb = O.solve_f64(nx, ny, wx, wy, b)
    pab = O.solve_f64(nx, ny, wx, wy, (a + b).astype(np.float32))
    for q in range(3):
        assert np.allclose(pab[q], pa[q] + pb[q], atol=1e-5)
    # the DC term is dropped: adding a constant changes nothing
    pc = O.solve_f64(nx, ny, wx, wy, (a + 5.0).astype(np.float32))
    for q in range(3):
        assert np.allclose(pc[q], pa[q], atol=1e-5)


_2D = {"orc_ddct2d_fwd": ("dct2", "dct2"), "orc_ddct2d_inv": ("dct3", "dct3"), "orc_ddsct2d_inv": ("dct3", "dst3"),
       "orc_ddcst2d_inv": ("dst3", "dct3")}


def _by_lines(fn2d, a, sfx):
    """The 2-D transform as the one-sequence transforms over every row, then every column."""
    rk, ck = _2D[fn2d]
    dt = np.float32 if sfx == "f32" else np.float64
    out = np.array(a, dt, copy=True)
    for i in range(out.shape[0]):
        out[i] = _run(f"orc_{rk}_{sfx}", out[i], dt)
    for j in range(out.shape[1]):
        out[:, j] = _run(f"orc_{ck}_{sfx}", out[:, j], dt)
    return out


@pytest.mark.parametrize("n1,n2", [(1, 8), (2, 4), (8, 8), (12, 16), (16, 12), (24, 48), (10, 16), (16, 10),
                                   (64, 32), (256, 512), (768, 96)])
@pytest.mark.parametrize("sfx", ["f32", "f64"])
@pytest.mark.parametrize("isa", [-1, 0])
def test_2d_transforms_are_the_1d_transforms_bit_for_bit(n1, n2, sfx, isa):
    """The 2-D transforms run eight rows (columns) at once, one in each lane of a vector: every
    lane must give the one-sequence transform's bits, with partial groups (12 and 10 rows or
    columns), the radix-3 lengths, lengths without a fast transform (10) and several threads,
    in the AVX2 build of the batched code (where the host has AVX2) and the baseline one."""
    dt = np.float32 if sfx == "f32" else np.float64
    x = np.random.default_rng(n1 * 1000 + n2).standard_normal((n1, n2)).astype(dt)
    for fn in _2D:
        want = _by_lines(fn, x, sfx)
        for threads in (1, 3):
            O.set_threads(threads)
            O.lib().orc_set_batch_isa(isa)
            try:
                got = np.array(x, copy=True)
                getattr(O.lib(), f"{fn}_{sfx}")(n1, n2, O.f32(got) if sfx == "f32" else O.f64(got))
            finally:
                O.set_threads(1)
                O.lib().orc_set_batch_isa(-1)
            assert np.array_equal(got.view(np.uint32 if sfx == "f32" else np.uint64),
                                  want.view(np.uint32 if sfx == "f32" else np.uint64)), (fn, threads)
