"""On-disk formats (hype_write_vec_binary / hype_read_vec_binary /
hype_write_vts; host C): PETSc binary Vec restart files
(post_process.c:47-63, hype.c:316-330) and VTK XML structured-grid snapshots
(post_process.c:29-41, read by plot.py:182-210)."""
import os
import re

import numpy as np
import pytest

import hype2d as H


def test_restart_layout_is_petsc_binary_vec(tmp_path):
    U = np.random.default_rng(1).standard_normal((5, 7, 6))
    path = tmp_path / "restart1.bin"
    H.write_restart(path, U)
    raw = path.read_bytes()
    assert len(raw) == 8 + 8 * U.size
    assert np.frombuffer(raw[:8], dtype=">i4").tolist() == [1211214, U.size]  # VEC_FILE_CLASSID, n
    assert np.array_equal(np.frombuffer(raw[8:], dtype=">f8"), U.ravel())  # [j][i][c], big-endian
    assert (tmp_path / "restart1.bin.info").read_text() == "-vecload_block_size 6\n"
    assert np.array_equal(H.read_restart(path, U.shape), U)


def test_restart_rejects_wrong_length_and_missing_file(tmp_path):
    U = np.ones((4, 4, 4))
    path = tmp_path / "r.bin"
    H.write_restart(path, U)
    with pytest.raises(H.HypeError) as e:
        H.read_restart(path, (4, 4, 6))
    assert e.value.code == H.E_INVALID
    with pytest.raises(H.HypeError) as e:
        H.read_restart(tmp_path / "absent.bin", (4, 4, 4))
    assert e.value.code == H.E_NOTFOUND


def parse_vts(path):
    raw = open(path, "rb").read()
    head, data = raw.split(b'<AppendedData encoding="raw">', 1)
    data = data[data.index(b"_") + 1:]
    head = head.decode()
    ext = [int(v) for v in re.search(r'WholeExtent="([^"]+)"', head).group(1).split()]
    arrays = {}
    for m in re.finditer(r'<DataArray type="Float64" Name="([^"]+)"(?: NumberOfComponents="(\d)")? '
                         r'format="appended" offset="(\d+)"/>', head):
        name, ncomp, off = m.group(1), int(m.group(2) or 1), int(m.group(3))
        nbytes = int(np.frombuffer(data[off:off + 8], dtype="<u8")[0])
        arrays[name] = np.frombuffer(data[off + 8:off + 8 + nbytes], dtype="<f8").reshape(-1, ncomp)
    assert 'byte_order="LittleEndian"' in head and 'header_type="UInt64"' in head
    return ext, arrays


@pytest.mark.parametrize("coord", [H.COORD_EXACT, H.COORD_DMDA])
def test_vts_snapshot_round_trip(tmp_path, coord):
    nx, ny = 9, 5
    p = H.make_problem(H.KAPILA, nx, ny, domain=(0.0, 1.8, 0.0, 1.0), bc=(H.BC_PERIODIC,) * 2 + (H.BC_REFLECTIVE,) * 2,
                       gamma1=1.4, gamma2=1.648)
    W = np.random.default_rng(2).uniform(0.5, 2.0, (ny, nx, 6))
    path = tmp_path / "sol-00000000.vts"
    H.write_vts(path, p, W, coord=coord)
    ext, arrays = parse_vts(path)
    assert ext == [0, nx - 1, 0, ny - 1, 0, 0]
    names = ["rho_1", "rho_2", "u", "v", "p", "alpha_1"]
    assert list(arrays)[1:] == names  # one scalar array per component, in component order
    for c, name in enumerate(names):
        assert np.array_equal(arrays[name][:, 0], W[..., c].ravel())  # i fastest, as plot.py reshapes
    h = 1.8 / nx
    pts = arrays["Position"].reshape(ny, nx, 3)
    x = 0.0 + 0.5 * h + np.arange(nx) * h  # periodic x: spacing h either way
    if coord == H.COORD_EXACT:
        y = 0.0 + 0.5 * h + np.arange(ny) * h
    else:  # DMDASetUniformCoordinates on a non-periodic axis: span over n - 1 gaps
        y = 0.0 + 0.5 * h + np.arange(ny) * (((1.0 + 0.5 * h) - (0.0 + 0.5 * h)) / (ny - 1))
    assert np.array_equal(pts[0, :, 0], x) and np.array_equal(pts[:, 0, 1], y) and not pts[..., 2].any()
