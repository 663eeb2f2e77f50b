"""Host logic of libhype2d.so that runs without a GPU: the tile-shape model,
the tile partition, the XCD-aware order, the error-word decoding and the
left-inlet table.  These are internal functions of csrc/host/common.c and
ic.c, reached through the library's exported symbols."""
import ctypes as C
import struct

import numpy as np
import pytest

import hype2d as H
import oracle_py as orc
import states


class Tile(C.Structure):
    _fields_ = [("inst", C.c_int), ("x0", C.c_int), ("y0", C.c_int), ("w", C.c_int), ("h", C.c_int),
                ("flags", C.c_int)]


@pytest.fixture(scope="module")
def L():
    lib = H.lib()
    lib.hype_tile_rows.argtypes = [C.c_int, C.c_int]
    lib.hype_tile_rows.restype = C.c_double
    lib.hype_tile_cost.argtypes = [C.c_long, C.c_double, C.c_int, C.c_int]
    lib.hype_tile_cost.restype = C.c_double
    lib.hype_pick_segments.argtypes = [C.c_int] * 6
    lib.hype_make_tiles.argtypes = [C.c_int] * 5 + [C.POINTER(Tile), C.c_int]
    lib.hype_xcd_order.argtypes = [C.POINTER(Tile), C.c_int, C.c_int]
    lib.hype_key_bits.argtypes = [C.c_int, C.c_int]
    lib.hype_decode_error.argtypes = [C.c_ulonglong, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                      C.c_double, C.c_char_p, C.c_size_t, C.POINTER(C.c_double),
                                      C.POINTER(C.c_double)]
    lib.hype_decode_error.restype = None
    lib.hype_inlet_doubles.argtypes = [C.POINTER(H.Problem)]
    lib.hype_inlet_doubles.restype = C.c_size_t
    lib.hype_inlet_table.argtypes = [C.POINTER(H.Problem), C.POINTER(C.c_double)]
    lib.hype_inlet_table.restype = None
    return lib


def test_tile_rows_and_cost(L):
    assert L.hype_tile_rows(9, 0) == 12.0  # h + two ring rows + prologue
    assert L.hype_tile_rows(18, 1) == 9 + 2.25  # half the rows, one ring row, the middle face
    assert L.hype_tile_rows(17, 1) == 9 + 2.25  # the lower half takes the odd row
    assert L.hype_tile_rows(1, 1) == 4.0  # a one-row tile is not split
    # waves beyond one per SIMD queue behind, with the two-wave latency gain when resident
    assert L.hype_tile_cost(1024, 10.0, 1024, 1) == 10.0
    assert L.hype_tile_cost(1025, 10.0, 1024, 1) == 20.0
    assert L.hype_tile_cost(2048, 10.0, 1024, 2) == pytest.approx(20.0 / 1.5)


@pytest.mark.parametrize("nx,ny,occ,pair,want", [
    (1024, 512, 1, 1, 29),   # Kapila 1024x512, split tiles: 493 two-wave tiles of 17-18 rows
    (1024, 512, 1, 0, 57),   # the same with one wave per tile: 969 tiles of 9 rows
    (256, 256, 2, 0, 128),   # Euler 256^2, one-wave tiles at two waves per SIMD: 2-row tiles
    (256, 256, 1, 1, 64),    # Euler 256^2, split tiles: 4-row tiles
])
def test_pick_segments_for_the_benchmark_grids(L, nx, ny, occ, pair, want):
    assert L.hype_pick_segments(nx, ny, 1024, occ, 0, pair) == want


def make_tiles(L, nx, ny, nsy):
    nsx = (nx + 61) // 62
    n = nsx * nsy
    arr = (Tile * n)()
    assert L.hype_make_tiles(0, nx, ny, nsx, nsy, arr, n) == n
    return arr, n


@pytest.mark.parametrize("nx,ny,nsy", [(1024, 512, 29), (130, 29, 5), (62, 3, 3), (63, 7, 2)])
def test_tiles_partition_the_grid(L, nx, ny, nsy):
    arr, n = make_tiles(L, nx, ny, nsy)
    cover = np.zeros((ny, nx), dtype=int)
    for t in arr:
        assert 1 <= t.w <= 62 and t.h >= 1
        cover[t.y0:t.y0 + t.h, t.x0:t.x0 + t.w] += 1
    assert np.all(cover == 1)
    hs = {t.h for t in arr}
    assert max(hs) - min(hs) <= 1  # balanced to within one row
    assert [t.flags for t in arr] == [1] + [0] * (n - 1)  # one leader per instance


def test_xcd_order_gives_each_xcd_a_contiguous_band(L):
    arr, n = make_tiles(L, 1024, 512, 29)
    orig = [(t.x0, t.y0) for t in arr]
    assert L.hype_xcd_order(arr, n, 8) == 0
    new = [(t.x0, t.y0) for t in arr]
    assert sorted(new) == sorted(orig)  # a permutation
    pos = {xy: k for k, xy in enumerate(orig)}
    start = 0
    for x in range(8):
        band = [pos[new[i]] for i in range(x, n, 8)]  # workgroups that run on XCD x
        assert band == list(range(start, start + len(band)))  # consecutive in row-major order
        start += len(band)
    assert start == n


def test_key_bits_threshold(L):
    assert L.hype_key_bits(1024, 512) == 28
    assert L.hype_key_bits(8200, 8200) == 36  # past 2^28 face points


def pack(seq, key, kind, val, K):
    f = struct.unpack("<I", struct.pack("<f", val))[0]
    return (seq << 60) | ((key & ((1 << K) - 1)) << (60 - K)) | (kind << (60 - K - 4)) | (f >> (K - 24))


@pytest.mark.parametrize("nx,ny", [(100, 40), (8200, 8200)])
def test_error_word_decodes_to_the_reference_message(L, nx, ny):
    K = L.hype_key_bits(nx, ny)
    ngf, h, x_min, y_min = 2, 0.01, -0.5, 0.25
    msg = C.create_string_buffer(256)
    x, y = C.c_double(), C.c_double()
    # an x-face point: face (i, j) of the (nx+1) x ny x-faces, point 1
    i, j = 37, 23
    key = (j * (nx + 1) + i) * ngf + 1
    L.hype_decode_error(pack(3, key, 2, -1.5, K), nx, ny, ngf, x_min, y_min, h, msg, 256, C.byref(x), C.byref(y))
    assert msg.value.decode() == f"Negative pressure, p = {-1.5:f}\nAt x = {x_min + i * h:f}, y = {y_min + j * h:f}"
    assert (x.value, y.value) == (x_min + i * h, y_min + j * h)
    # a y-face point: after all x-face points, face (i, j) of the nx x (ny+1) y-faces
    i, j = 5, 40
    key = ny * (nx + 1) * ngf + (j * nx + i) * ngf
    L.hype_decode_error(pack(1, key, 3, -0.25, K), nx, ny, ngf, x_min, y_min, h, msg, 256, C.byref(x), C.byref(y))
    assert msg.value.decode().startswith(f"Negative square speed of sound = {-0.25:f}")
    assert (x.value, y.value) == (x_min + i * h, y_min + j * h)


@pytest.mark.parametrize("prim", [True, False])
def test_inlet_table_equals_the_oracle_profile(L, prim):
    """The host table of the Kapila left inlet (one row per ghosted row j =
    -3 .. ny+2) is the oracle's InletBCPrim state, or its PDEPrim2Cons, bit
    for bit."""
    nx, ny = 30, 24
    h = 3.0 / ny
    eos = states.INLET_EOS
    p = H.make_problem(H.KAPILA, nx, ny, order=2, prim=prim, domain=(0.0, nx * h, -1.5, 1.5),
                       bc=(H.BC_INLET, H.BC_TRANSMISSIVE, H.BC_REFLECTIVE, H.BC_REFLECTIVE), **eos)
    op = orc.problem(orc.KAPILA, nx, ny, N=2, prim=prim, x_min=0.0, y_min=-1.5, h=h,
                     bc=(orc.BC_INLET, orc.BC_TRANSMISSIVE, orc.BC_REFLECTIVE, orc.BC_REFLECTIVE),
                     g1=eos["gamma1"], g2=eos["gamma2"], p1=eos["pi1"], p2=eos["pi2"])
    n = L.hype_inlet_doubles(C.byref(p))
    assert n == (ny + 6) * 6
    tab = np.zeros(n)
    L.hype_inlet_table(C.byref(p), tab.ctypes.data_as(C.POINTER(C.c_double)))
    tab = tab.reshape(ny + 6, 6)
    for j in range(-3, ny + 3):
        V = orc.inlet_prim(op, -1.5 + (j + 0.5) * h)
        want = V if prim else orc.prim2cons(op, V)
        assert np.array_equal(tab[j + 3], want), j
    p.bc = (H.BC_TRANSMISSIVE,) * 4
    assert L.hype_inlet_doubles(C.byref(p)) == 0


def test_served_stream_prefixes_are_prefixes():
    """Every request's draws follow the previous request's, so the 16-, 64-,
    256- and 1,024-request traces the tests and benchmark replay are the
    first requests of the one 2,400-request stream (20 requests/s for
    120 s), with the stated mix and ranges."""
    import workloads as WL
    full = WL.served_draws(np.random.default_rng(2024), 2400, 20.0)
    for n in (16, 64, 256, 1024):
        assert WL.served_draws(np.random.default_rng(2024), n, 20.0) == full[:n]
    t = np.array([d[0] for d in full])
    assert np.all(np.diff(t) > 0) and 100.0 < t[-1] < 140.0
    sides = np.array([d[2] for d in full])
    steps = np.array([d[3] for d in full])
    assert sides.min() >= 64 and sides.max() < 2048 and steps.min() >= 100 and steps.max() <= 2000
    share = [np.mean([d[1] == s for d in full]) for s in (H.EULER, H.KAPILA, H.BAER_NUNZIATO)]
    assert abs(share[0] - 0.50) < 0.04 and abs(share[1] - 0.35) < 0.04 and abs(share[2] - 0.15) < 0.03
    # a BN request carries its own interface seed, the others none
    assert all((d[4] != 0) == (d[1] == H.BAER_NUNZIATO) for d in full)
