"""Files at the boundary: the geometry reader (readGeometry, src/complab_functions.hh:127-167)
and the VTK ImageData writer (writeAdvVTI / writeNsVTI / saveGeometry,
src/complab_functions.hh:86-125), in Python (clb.read_geometry / clb.write_vti) and in the
C++ host header (complab_gpu::readGeometry / writeVTI, through tests/cpp/io_check).

The reference's reader consumes the file slice by slice along x, while its makeGeometry.py
(examples/makeGeometry.py:5-7, 74-80) writes one integer per line with x fastest; both orders
are read and checked against the voxel each value names. The VTI files are parsed back
(XML + base64, UInt32 length header) and compared with the fields written."""
import base64
import os
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import clb

HERE = os.path.dirname(os.path.abspath(__file__))
NX, NY, NZ = 7, 5, 4


def coded_geometry():
    """A geometry whose value tells its voxel: code = x + 10 y + 50 z (fits a byte)."""
    z, y, x = np.meshgrid(np.arange(NZ), np.arange(NY), np.arange(NX), indexing="ij")
    return (x + 10 * y + 50 * z).astype(np.int64)  # (nz, ny, nx)


def write_dat(path, g, order):
    with open(path, "w") as f:
        if order == clb.GEOM_X_FASTEST:
            for z in range(NZ):
                for y in range(NY):
                    for x in range(NX):
                        f.write("%d\n" % g[z, y, x])
        else:
            for x in range(NX):
                for y in range(NY):
                    for z in range(NZ):
                        f.write("%d\n" % g[z, y, x])


def expected_mask(g):
    out = np.empty((NZ, NY, NX + 2), dtype=np.uint8)
    out[:, :, 1:-1] = g
    out[:, :, 0] = g[:, :, 0]
    out[:, :, -1] = g[:, :, -1]
    return out.ravel()


def read_vti(path):
    root = ET.parse(path).getroot()
    assert root.get("type") == "ImageData" and root.get("byte_order") == "LittleEndian"
    ext = [int(v) for v in root.find("ImageData").get("WholeExtent").split()]
    types = {"Float64": "<f8", "Float32": "<f4", "Int32": "<i4", "UInt8": "u1", "Int64": "<i8"}
    arrays = {}
    for da in root.iter("DataArray"):
        raw = base64.b64decode(da.text.strip())
        n = int(np.frombuffer(raw[:4], "<u4")[0])
        a = np.frombuffer(raw[4:4 + n], types[da.get("type")])
        arrays[da.get("Name")] = a.reshape(-1, int(da.get("NumberOfComponents")))
    return ext, arrays


@pytest.mark.parametrize("order", [clb.GEOM_X_FASTEST, clb.GEOM_X_SLICES])
def test_read_geometry_orders(tmp_path, order):
    g = coded_geometry()
    p = tmp_path / "geom.dat"
    write_dat(p, g, order)
    m = clb.read_geometry(str(p), NX, NY, NZ, order)
    assert np.array_equal(m, expected_mask(g))
    # the other order reads a different lattice (the orders are not interchangeable)
    other = clb.GEOM_X_SLICES if order == clb.GEOM_X_FASTEST else clb.GEOM_X_FASTEST
    assert not np.array_equal(clb.read_geometry(str(p), NX, NY, NZ, other), m)


def test_read_geometry_rejects(tmp_path):
    p = tmp_path / "short.dat"
    p.write_text("1\n2\n3\n")
    with pytest.raises(clb.ClbError):
        clb.read_geometry(str(p), NX, NY, NZ)
    p.write_text("\n".join(["300"] * (NX * NY * NZ)))
    with pytest.raises(clb.ClbError):
        clb.read_geometry(str(p), NX, NY, NZ)


def test_write_vti_round_trip(tmp_path):
    nx, ny, nz = NX + 2, NY, NZ
    N = nx * ny * nz
    rng = np.random.default_rng(3)
    dens = rng.random(N)
    vel = rng.random((N, 3)).astype(np.float32)
    tag = rng.integers(0, 4, N).astype(np.uint8)
    p = tmp_path / "f.vti"
    clb.write_vti(str(p), nx, ny, nz, {"Density": dens, "velocity": vel, "tag": tag})
    ext, arr = read_vti(str(p))
    assert ext == [0, nx - 3, 0, ny - 1, 0, nz - 1]
    crop = lambda a, c: np.asarray(a).reshape(nz, ny, nx, c)[:, :, 1:nx - 1, :].reshape(-1, c)  # noqa: E731
    assert np.array_equal(arr["Density"], crop(dens, 1))
    assert arr["velocity"].dtype == np.float32 and np.array_equal(arr["velocity"], crop(vel, 3))
    assert np.array_equal(arr["tag"], crop(tag, 1))


@pytest.mark.parametrize("order", [0, 1])
def test_cpp_read_geometry_and_vti(tmp_path, order):
    exe = os.path.join(HERE, "cpp", "io_check")
    if not os.path.exists(exe):
        subprocess.run(["make", "-C", os.path.join(HERE, "cpp"), "io_check"], check=True, capture_output=True)
    g = coded_geometry()
    p = tmp_path / "geom.dat"
    write_dat(p, g, clb.GEOM_X_SLICES if order else clb.GEOM_X_FASTEST)
    mb, vti = tmp_path / "mask.bin", tmp_path / "out.vti"
    r = subprocess.run([exe, str(p), str(NX), str(NY), str(NZ), str(order), str(mb), str(vti)], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    m = np.fromfile(mb, dtype=np.uint8)
    assert np.array_equal(m, expected_mask(g))
    ext, arr = read_vti(str(vti))
    assert ext == [0, NX - 1, 0, NY - 1, 0, NZ - 1]
    assert np.array_equal(arr["tag"][:, 0], g.ravel().astype(np.float64))  # the cropped planes are the file's
    z, y, x = np.meshgrid(np.arange(NZ), np.arange(NY), np.arange(1, NX + 1), indexing="ij")
    assert np.array_equal(arr["xyz"], np.stack([x.ravel(), y.ravel(), z.ravel()], 1).astype(np.float64))
    # a missing file is an error, not a crash
    r = subprocess.run([exe, str(tmp_path / "none.dat"), "3", "3", "3", "0", str(mb), str(vti)], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 3


# ---- the reference's own geometry files (examples/input_shared, copied unchanged under
# tests/golden/) at the size their CompLaB.xml gives: nx = ny = 24, nz = 6 (e.g.
# examples/08_two_microbes/CompLaB.xml:39-52, material numbers pore 2, solid 0, bounce-back
# 1, microbe0 3, microbe1 4). The checks of examples/validateExamples.py:150-178 (value count,
# every microbe code present), then the layout examples/makeGeometry.py:33-109 builds, which
# fixes the file order: bounce-back rows at y = 0 and y = 23, the patches, every z plane alike.
GOLDEN = os.path.join(HERE, "golden")
REF_DECLARED = {0, 1, 2, 3, 4}


def ref_layout(kind):
    g = np.full((6, 24, 24), 2, dtype=np.int64)
    g[:, 0, :] = 1
    g[:, 23, :] = 1
    if kind == "two_microbes":
        g[:, 1:4, 4:10] = 3
        g[:, 20:23, 14:20] = 4
    else:
        yy, xx = np.mgrid[0:24, 0:24]
        for cx, cy in ((7, 8), (7, 16), (16, 8), (16, 16)):
            g[:, (xx - cx) ** 2 + (yy - cy) ** 2 <= 9] = 0
    return g


@pytest.mark.parametrize("kind", ["two_microbes", "grains"])
def test_reference_geometry_files(kind):
    path = os.path.join(GOLDEN, f"{kind}_24x24x6.dat")
    vals = [int(v) for v in open(path).read().split()]
    assert len(vals) == 24 * 24 * 6
    assert set(vals) <= REF_DECLARED
    if kind == "two_microbes":
        assert {3, 4} <= set(vals)                     # every microbe is seeded somewhere
    m = clb.read_geometry(path, 24, 24, 6, clb.GEOM_X_FASTEST).reshape(6, 24, 26)
    g = m[:, :, 1:-1]
    assert np.array_equal(g, ref_layout(kind))
    assert np.array_equal(m[:, :, 0], g[:, :, 0]) and np.array_equal(m[:, :, -1], g[:, :, -1])
    assert all(np.array_equal(g[z], g[0]) for z in range(6))
    counts = np.bincount(np.asarray(vals), minlength=5)
    if kind == "two_microbes":
        assert counts.tolist() == [0, 288, 3456 - 288 - 216, 108, 108]
    # the C++ reader gives the same lattice mask
    exe = os.path.join(HERE, "cpp", "io_check")
    if not os.path.exists(exe):
        subprocess.run(["make", "-C", os.path.join(HERE, "cpp"), "io_check"], check=True, capture_output=True)
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        mb, vti = os.path.join(td, "mask.bin"), os.path.join(td, "out.vti")
        r = subprocess.run([exe, path, "24", "24", "6", "0", mb, vti], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, r.stderr
        assert np.array_equal(np.fromfile(mb, dtype=np.uint8), m.ravel())
