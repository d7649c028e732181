"""wavesim's coarse-to-fine Sampler (crates/wavesim/src/lib.rs:1548-1616) on the oracle, pinned by
what bilinear interpolation fixes: a surface and velocities linear in x and y are reproduced
exactly wherever all four corners are wet; dry corners drop out and the rest are renormalised;
no wet corner gives no water; cells beyond the coarse grid use its edge (one ghost layer)."""

from __future__ import annotations

import numpy as np

from oracle.oracle import OracleSampler
from wavesim_amd.grid import Grid, State


def coarse_field(grid, eta_fn, u_fn, v_fn, bed_fn):
    s = State.zeros(grid)
    w = grid.width()
    b = np.zeros(grid.cells())
    for j in range(grid.ny + 6):
        for i in range(w):
            x = (i - 3 + 0.5) * grid.dx
            y = (j - 3 + 0.5) * grid.dy
            k = j * w + i
            b[k] = bed_fn(x, y)
    inner = grid.inner
    X = (np.arange(grid.nx) + 0.5) * grid.dx
    Y = (np.arange(grid.ny) + 0.5) * grid.dy
    XX, YY = np.meshgrid(X, Y)
    bi = grid.inner(b)
    h = np.maximum(eta_fn(XX, YY) - bi, 0.0)
    inner(s.h)[:, :] = h
    inner(s.hu)[:, :] = h * u_fn(XX, YY)
    inner(s.hv)[:, :] = h * v_fn(XX, YY)
    s.fill_walls(grid)
    return s, b


def test_linear_water_is_sampled_exactly():
    cg = Grid(20, 12, 6.0, 6.0)
    eta = lambda x, y: 0.3 + 0.01 * x - 0.02 * y
    u = lambda x, y: 0.5 - 0.003 * x + 0.001 * y
    v = lambda x, y: -0.2 + 0.002 * x
    s, b = coarse_field(cg, eta, u, v, lambda x, y: -10.0 + 0.0 * x)
    # fine cells of 2 m, zone columns 0-5 of a fine grid starting 12 fine cells into the bed
    cells = [(i, j) for j in range(3, 3 + 20) for i in range(3, 9)]
    sm = OracleSampler(cells, 12, 2.0, 2.0, cg.nx, cg.ny, 6.0, 6.0, crop=1)
    out = sm.sample(s.h, s.hu, s.hv, b)
    for (i, j), (e, uu, vv) in zip(cells, out):
        x = (12 + i - 3 + 0.5) * 2.0  # in the bed file's frame
        y = (j - 3 + 0.5) * 2.0
        xc = x - 1 * 6.0  # the coarse grid starts one coarse column in (crop)
        if 3.0 <= xc <= cg.nx * 6.0 - 3.0 and 3.0 <= y <= cg.ny * 6.0 - 3.0:
            assert abs(e - eta(xc, y)) < 1e-12 and abs(uu - u(xc, y)) < 1e-12
            assert abs(vv - v(xc, y)) < 1e-12


def test_dry_corners_drop_out_and_no_water_is_minus_infinity():
    cg = Grid(8, 8, 4.0, 4.0)
    # the bed rises above the surface east of x = 16 m: columns 4.. are dry
    s, b = coarse_field(cg, lambda x, y: 0.0 * x, lambda x, y: 0.1 + 0.0 * x,
                        lambda x, y: 0.0 * x, lambda x, y: np.where(x < 16.0, -2.0, 1.0))
    cells = [(3 + 6, 5), (3 + 7, 5), (3 + 14, 5)]  # fine 2 m cells: x = 13, 15, 29 m
    sm = OracleSampler(cells, 0, 2.0, 2.0, cg.nx, cg.ny, 4.0, 4.0)
    out = sm.sample(s.h, s.hu, s.hv, b)
    assert out[0][0] == 0.0 and abs(out[0][1] - 0.1) < 1e-15  # all four wet
    assert out[1][0] == 0.0 and abs(out[1][1] - 0.1) < 1e-15  # two wet corners, renormalised
    assert out[2][0] == float("-inf") and out[2][1] == 0.0 and out[2][2] == 0.0


def test_cells_beyond_the_coarse_grid_use_its_edge():
    cg = Grid(5, 4, 10.0, 10.0)
    s, b = coarse_field(cg, lambda x, y: 1.0 + 0.1 * x, lambda x, y: 0.0 * x,
                        lambda x, y: 0.0 * x, lambda x, y: -5.0 + 0.0 * x)
    # a fine cell far east of the coarse grid: every corner clamps to the first ghost column,
    # which mirrors the last interior column
    sm = OracleSampler([(3 + 40, 4)], 0, 5.0, 5.0, cg.nx, cg.ny, 10.0, 10.0)
    e = sm.sample(s.h, s.hu, s.hv, b)[0][0]
    assert e == 1.0 + 0.1 * 45.0


def test_lerp_is_linear_where_both_are_water_else_the_nearer():
    frm = np.array([[1.0, 2.0, 3.0], [float("-inf"), 0.0, 0.0]])
    to = np.array([[3.0, 2.0, 1.0], [0.5, 1.0, 1.0]])
    mid = OracleSampler.lerp(frm, to, 0.25)
    assert np.array_equal(mid[0], [1.5, 2.0, 2.5])
    assert np.array_equal(mid[1], frm[1])
    assert np.array_equal(OracleSampler.lerp(frm, to, 0.5)[1], to[1])
