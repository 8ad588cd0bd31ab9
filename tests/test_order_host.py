"""Host logic of the state order: capi.brick_order against a direct loop
over bricks, its permutation property, cut bricks at the high faces, and
the locality it buys (the spread of a brick's unknowns in the caller's
numbering against the spread of its slots in the internal one)."""
import numpy as np
import pytest

from mcls_amd import capi


def _loop_order(nx, ny, nz, bx, by, bz):
    out = []
    for gz in range(0, nz, bz):
        for gy in range(0, ny, by):
            for gx in range(0, nx, bx):
                for z in range(gz, min(gz + bz, nz)):
                    for y in range(gy, min(gy + by, ny)):
                        for x in range(gx, min(gx + bx, nx)):
                            out.append(x + nx * (y + ny * z))
    return np.array(out, np.int32)


@pytest.mark.parametrize("dims", [(8, 8, 8, 4, 4, 4), (7, 5, 6, 3, 2, 4), (9, 4, 3, 9, 4, 3), (5, 5, 5, 1, 1, 1),
                                  (6, 7, 8, 16, 16, 16)])
def test_brick_order_matches_loop(dims):
    o = capi.brick_order(*dims)
    assert o.dtype == np.int32
    assert np.array_equal(o, _loop_order(*dims))
    n = dims[0] * dims[1] * dims[2]
    assert np.array_equal(np.sort(o), np.arange(n))


def test_brick_order_is_natural_for_whole_grid_bricks():
    assert np.array_equal(capi.brick_order(6, 5, 4, 6, 5, 4), np.arange(120))


def test_brick_order_neighbours_stay_close():
    nx, b = 32, 8
    o = capi.brick_order(nx, nx, nx, b, b, b)
    inv = np.empty_like(o)
    inv[o] = np.arange(len(o))
    # internal slots of the +z neighbours inside a brick are b*b apart
    # instead of nx*nx in the caller's numbering
    z_in = [(inv[u + nx * nx] - inv[u]) for u in range(nx * nx * (nx - 1)) if (u // (nx * nx)) % b != b - 1]
    assert set(z_in) == {b * b}


def test_brick_order_rejects_bad_sizes():
    with pytest.raises(ValueError):
        capi.brick_order(4, 4, 0, 2, 2, 2)
    with pytest.raises(ValueError):
        capi.brick_order(4, 4, 4, 2, 0, 2)


def _nested_loop(nx, ny, nz, levels):
    """Recursive definition: bricks of levels[0] x-fastest, each visited by
    the remaining levels, the points of the last level x-fastest."""
    out = []

    def visit(x0, y0, z0, x1, y1, z1, lv):
        if lv == len(levels):
            for z in range(z0, z1):
                for y in range(y0, y1):
                    for x in range(x0, x1):
                        out.append(x + nx * (y + ny * z))
            return
        bx, by, bz = levels[lv]
        for gz in range(z0, z1, bz):
            for gy in range(y0, y1, by):
                for gx in range(x0, x1, bx):
                    visit(gx, gy, gz, min(gx + bx, x1), min(gy + by, y1), min(gz + bz, z1), lv + 1)

    visit(0, 0, 0, nx, ny, nz, 0)
    return np.array(out, np.int32)


@pytest.mark.parametrize("dims,levels", [((16, 16, 16), [(8, 8, 8), (4, 4, 4)]),
                                         ((12, 8, 10), [(6, 4, 5), (3, 2, 5), (1, 1, 1)]),
                                         ((9, 7, 5), [(4, 4, 4), (2, 2, 2)]),
                                         ((8, 8, 8), [(4, 2, 8)])])
def test_nested_brick_order_matches_recursion(dims, levels):
    o = capi.nested_brick_order(*dims, levels)
    assert np.array_equal(o, _nested_loop(*dims, levels))
    assert np.array_equal(np.sort(o), np.arange(dims[0] * dims[1] * dims[2]))


def test_nested_brick_order_one_level_is_brick_order():
    assert np.array_equal(capi.nested_brick_order(10, 9, 8, [(4, 3, 2)]), capi.brick_order(10, 9, 8, 4, 3, 2))


def test_nested_brick_order_rejects_non_nesting_sizes():
    with pytest.raises(ValueError):
        capi.nested_brick_order(8, 8, 8, [(4, 4, 4), (3, 3, 3)])
    with pytest.raises(ValueError):
        capi.nested_brick_order(8, 8, 8, [])
