"""The reference's file formats through the C ABI (host code, no GPU):
PETSc binary Vec files (VecView / VecLoadIntoVector of PETSc 3.1, used for the
vfield/ufield/pfield/nvfield checkpoints of main.c:376-934) and the grid
readers of init.c:258-394 (grid.dat text and binary, xyz.dat).

The Vec format is pinned byte by byte against PETSc's published layout
(big-endian int32 cookie 1211214, int32 length, big-endian doubles), written
here with the struct module; the xyz.dat reader against the reference's own
channel deck (tests/golden/channel_retau3000/xyz.dat), read independently with
numpy; the grid.dat readers against files written from known coordinates.
"""
import struct

import numpy as np
import pytest

from helpers import cases, vfs

GOLDEN = __import__("pathlib").Path(__file__).resolve().parent / "golden"


def test_vec_write_layout(tmp_path):
    rng = np.random.default_rng(1)
    v = rng.standard_normal(1000)
    v[3] = -0.0
    v[7] = np.inf
    f = tmp_path / "vfield000010_0.dat"
    vfs.vec_write(f, v)
    raw = f.read_bytes()
    assert raw[:8] == struct.pack(">ii", 1211214, 1000)
    assert raw[8:] == v.astype(">f8").tobytes()
    assert len(raw) == 8 + 8 * 1000


def test_vec_read_petsc_file(tmp_path):
    rng = np.random.default_rng(2)
    v = rng.standard_normal((6, 5, 4, 3))
    f = tmp_path / "ufield000003_0.dat"
    f.write_bytes(struct.pack(">ii", 1211214, v.size) + v.astype(">f8").tobytes())
    assert vfs.vec_length(f) == v.size
    got = vfs.vec_read(f, v.shape)
    assert np.array_equal(got, v) and got.dtype == np.float64


def test_vec_read_rejects_bad_files(tmp_path):
    f = tmp_path / "x.dat"
    f.write_bytes(struct.pack(">ii", 1211214, 10) + np.zeros(10, ">f8").tobytes())
    with pytest.raises(vfs.VFSError, match="length 10 in the file, 12 expected"):
        vfs.vec_read(f, (12,))
    f.write_bytes(struct.pack(">ii", 1211215, 10) + np.zeros(10, ">f8").tobytes())
    with pytest.raises(vfs.VFSError, match="class cookie"):
        vfs.vec_read(f)
    f.write_bytes(struct.pack(">ii", 1211214, 10) + np.zeros(9, ">f8").tobytes())
    with pytest.raises(vfs.VFSError, match="short"):
        vfs.vec_read(f, (10,))
    with pytest.raises(vfs.VFSError, match="cannot open"):
        vfs.vec_read(tmp_path / "missing.dat", (10,))


def test_vec_round_trip_empty(tmp_path):
    f = tmp_path / "e.dat"
    vfs.vec_write(f, np.zeros(0))
    assert vfs.vec_length(f) == 0 and vfs.vec_read(f).size == 0


def test_xyz_reader_channel_deck():
    path = GOLDEN / "channel_retau3000" / "xyz.dat"
    im, jm, km = vfs.grid_dims(path, "xyz")
    assert (im, jm, km) == (121, 41, 61)
    rows = np.loadtxt(path, skiprows=1)
    assert rows.shape == (im + jm + km, 3)
    X, Y, Z = rows[:im, 0], rows[im:im + jm, 1], rows[im + jm:, 2]
    coor = vfs.grid_read(path, "xyz", cl=2.0, L_dim=3.0)
    assert coor.shape == (km + 1, jm + 1, im + 1, 3)
    # coor[k][j][i] = (X[i], Y[j], Z[k]) / cl * L_dim; the last DA plane stays zero
    assert np.array_equal(coor[:km, :jm, :im, 0], np.broadcast_to(X / 2.0 * 3.0, (km, jm, im)))
    assert np.array_equal(coor[:km, :jm, :im, 1], np.broadcast_to((Y / 2.0 * 3.0)[:, None], (km, jm, im)))
    assert np.array_equal(coor[:km, :jm, :im, 2], np.broadcast_to((Z / 2.0 * 3.0)[:, None, None], (km, jm, im)))
    assert not coor[km].any() and not coor[:, jm].any() and not coor[:, :, im].any()
    # the deck's wall-normal grid is stretched and spans the half channel
    assert Y[0] == 0.0 and np.all(np.diff(Y) > 0) and np.all(np.diff(X) > 0)


def _grid_blocks():
    c1 = cases.channel(6, 4, 5)
    c2 = cases.channel(4, 6, 3, bump=0.05)
    blocks = []
    for c in (c1, c2):
        nx, ny, nz = c.cells
        blocks.append(c.coor[:nz + 1, :ny + 1, :nx + 1])  # node coordinates (KM, JM, IM, 3)
    return [c1, c2], blocks


def test_grid_dat_text_and_binary(tmp_path):
    cs, blocks = _grid_blocks()
    # text grid.dat: block count, then per block "IM JM KM" and all x, all y, all z in (k, j, i) order
    lines = [f"{len(blocks)}"]
    for b in blocks:
        km, jm, im = b.shape[:3]
        lines.append(f"{im} {jm} {km}")
        for comp in range(3):
            lines.extend(f"{v:.17e}" for v in b[..., comp].ravel())
    (tmp_path / "grid.dat").write_text("\n".join(lines) + "\n")
    # binary grid.dat (-binary_input): native int block count and sizes, native doubles
    with open(tmp_path / "gridb.dat", "wb") as f:
        f.write(np.int32(len(blocks)).tobytes())
        for b in blocks:
            km, jm, im = b.shape[:3]
            f.write(np.array([im, jm, km], np.int32).tobytes())
            for comp in range(3):
                f.write(np.ascontiguousarray(b[..., comp]).astype(np.float64).tobytes())
    for fmt, name in (("text", "grid.dat"), ("binary", "gridb.dat")):
        for bi, (c, b) in enumerate(zip(cs, blocks)):
            km, jm, im = b.shape[:3]
            assert vfs.grid_dims(tmp_path / name, fmt, bi) == (im, jm, km)
            coor = vfs.grid_read(tmp_path / name, fmt, bi)
            assert coor.shape == c.coor.shape
            assert np.array_equal(coor, c.coor), (fmt, bi)  # %.17e round-trips doubles exactly
        with pytest.raises(vfs.VFSError, match="no such block"):
            vfs.grid_dims(tmp_path / name, fmt, 2)


def test_grid_read_rejects_short_files(tmp_path):
    (tmp_path / "g.dat").write_text("1\n3 3 3\n" + "0.5\n" * 20)
    with pytest.raises(vfs.VFSError, match="short grid file"):
        vfs.grid_read(tmp_path / "g.dat", "text")
