"""On-disk formats (emil_code/Utils.py, tomo3D.py read/write_file, the
example drivers' run state and derrEvol.txt log).  Host-only.

The reference's Utils.py is Python 2 and imports tomo3D, which does not
parse under Python 3, so it cannot be run here; the MGH layout is pinned by
building files byte by byte from the format (512-byte ``KEY= value;``
header, Fortran-ordered data) and by the reference's own state files
(examples/data/sphericalRadon/{iternum,parms,derrEvol}.txt, committed under
tests/golden/state_sphericalRadon).
"""
import os

import numpy as np
import pytest

from ct2 import Utils

GOLD = os.path.join(os.path.dirname(__file__), "golden", "state_sphericalRadon")


def _mgh_bytes(items, data_f_order: bytes):
    hdr = "{\n" + "".join(f"{k} {v};\n" for k, v in items) + "}\n\n"
    return hdr.encode() + b" " * (512 - len(hdr)) + data_f_order


@pytest.mark.parametrize("dtype,name", [(np.float32, "float"), (np.float64, "double"),
                                        (np.int16, "short"), (np.uint16, "unsigned_short")])
def test_mgh_read_hand_built(tmp_path, dtype, name):
    a = (np.arange(4 * 5 * 3) % 17).astype(dtype).reshape(4, 5, 3)
    items = [("HEADER_BYTES=", 512), ("DIM=", 3), ("SIZE1=", 4), ("SIZE2=", 5),
             ("BYTE_ORDER=", "little_endian"), ("TYPE=", name), ("SIZE3=", 3)]
    p = tmp_path / "a.mgh"
    p.write_bytes(_mgh_bytes(items, a.astype(np.dtype(dtype).newbyteorder("<")).tobytes(order="F")))
    b = Utils.MGHread(str(p))
    assert b.dtype == dtype and b.shape == (4, 5, 3) and b.flags.c_contiguous
    np.testing.assert_array_equal(b, a)
    # 2-D and big-endian data
    a2 = a[:, :, 0]
    items2 = [("HEADER_BYTES=", 512), ("DIM=", 2), ("SIZE1=", 4), ("SIZE2=", 5),
              ("BYTE_ORDER=", "big_endian"), ("TYPE=", name)]
    p.write_bytes(_mgh_bytes(items2, a2.astype(np.dtype(dtype).newbyteorder(">")).tobytes(order="F")))
    np.testing.assert_array_equal(Utils.MGHread(str(p)), a2)


def test_mgh_write_layout_and_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    a = rng.random((6, 7, 5))                      # float64 is written as float32
    p = str(tmp_path / "w.mgh")
    Utils.MGHwrite(p, a, {"comment": "hello"})
    raw = open(p, "rb").read()
    assert len(raw) == 512 + a.size * 4
    hdr = raw[:512].decode()
    assert hdr.startswith("{\nHEADER_BYTES= 512;\nDIM= 3;\nSIZE1= 6;\nSIZE2= 7;\n"
                          "BYTE_ORDER= little_endian;\nTYPE= float;\nSIZE3= 5;\ncomment hello;\n}\n\n")
    assert hdr.rstrip(" ").endswith("}\n\n")
    assert raw[512:] == a.astype("<f4").tobytes(order="F")
    np.testing.assert_array_equal(Utils.MGHread(p), a.astype(np.float32))
    assert Utils.MGHheader(p)["comment"] == "hello"
    # int64 -> int16, write_type
    Utils.MGHwrite(p, np.arange(12).reshape(3, 4))
    assert Utils.MGHread(p).dtype == np.int16
    Utils.MGHwrite(p, np.arange(12).reshape(3, 4), write_type="uint16")
    assert Utils.MGHheader(p)["TYPE="] == "unsigned_short"
    with pytest.raises(ValueError):
        Utils.MGHwrite(p, a, {"x" * 40: "y" * 500})


def test_write_sino_and_image_headers(tmp_path):
    from ct2.tomo3D import image3D, sinogram3D
    s = sinogram3D("circular_conebeam", {"source_to_detector": 10., "radius": 5.}, (4, 6, 5))
    s.mat[:] = np.arange(s.mat.size, dtype=np.float32).reshape(s.mat.shape)
    p = str(tmp_path / "s.mgh")
    Utils.WriteSino(p, s, "a sinogram")
    h = Utils.MGHheader(p)
    assert h["img_type"] == "sinogram3D" and h["config_name"] == "circular_conebeam"
    assert h["shape"] == "(4, 6, 5)" and h["comment"] == "a sinogram"
    assert float(h["radius"]) == 5.0 and int(h["s_include_endpoints"]) == 1
    np.testing.assert_array_equal(Utils.MGHread(p), s.mat)
    im = image3D((5, 4, 3), 2.0, 3.0, 4.0, -1.0, -1.5, -2.0)
    im.mat[:] = 1.5
    Utils.WriteImage(p, im)
    h = Utils.MGHheader(p)
    assert h["img_type"] == "image3D" and float(h["zlen"]) == 4.0 and float(h["y0"]) == -1.5


@pytest.mark.parametrize("lat", ["R", "L"])
def test_ge_data_to_sino(tmp_path, lat):
    from ct2 import configs3D
    rng = np.random.default_rng(1)
    stack = rng.standard_normal((16, 12, 11)).astype(np.float32)   # SIZE1, SIZE2, views
    p = str(tmp_path / "ge.mgh")
    Utils.MGHwrite(p, stack)
    s = Utils.GEdata2sino(p, pixsize=0.1, Laterality=lat)
    want = stack.transpose(2, 0, 1)
    if lat == "L":
        want = want[:, ::-1, :]
    want = np.where(want < 0, 0, want)
    np.testing.assert_array_equal(s.mat, want)
    assert (s.ns, s.nu, s.nv) == (11, 16, 12) and s.u0 == 0.0 and s.v0 == pytest.approx(-0.6)
    ang = np.array(Utils.GE_ANGLES_R if lat == "R" else Utils.GE_ANGLES_L)
    f = configs3D.configs["tomosynthesis"]
    for ip in range(11):
        np.testing.assert_allclose(s.frame_vectors[ip],
                                   np.asarray(f((ang[ip] + 90) / 180 * np.pi,
                                                {"center_to_detector": 23.7, "radius": 44.3}),
                                              np.float32))
    with pytest.raises(ValueError):
        Utils.GEdata2sino(p)


def test_reference_state_files():
    """The reference's own run state reads back; its resume procedure (three
    read() calls on one handle, jawPhantomCCB.py:160-162) fails on it."""
    it, gf, reg_fac, meps = Utils.read_state(GOLD)
    assert (it, gf, reg_fac, meps) == (14271, 0.000153901062215, 1.0, -1.0)
    with open(os.path.join(GOLD, "parms.txt")) as f:
        first = f.read()
        with pytest.raises(ValueError):
            float(first)
        assert f.read() == ""


def test_state_round_trip_and_derr_log(tmp_path):
    d = str(tmp_path)
    vol = np.random.default_rng(2).random((4, 5, 6)).astype(np.float32)
    Utils.write_state(d, 12, np.float64(0.125), 0.99, -1.0, vol)
    assert Utils.read_state(d) == (12, 0.125, 0.99, -1.0)
    assert open(os.path.join(d, "parms.txt")).read() == "0.125\n0.99\n-1.0\n"
    np.testing.assert_array_equal(Utils.read_dat(os.path.join(d, "imageTemp.dat"), (4, 5, 6)), vol)
    ref = Utils.read_derr(os.path.join(GOLD, "derrEvol_head.txt"))
    assert ref.shape == (40, 4)
    np.testing.assert_array_equal(ref[0], [1, 171.505096436, 1359.05126953, 10.7059497833])
    for r in ref[:3]:
        Utils.append_derr(d, int(r[0]), np.float32(r[1]), r[2], r[3])
    back = Utils.read_derr(os.path.join(d, "derrEvol.txt"))
    np.testing.assert_array_equal(back[:, [0, 2, 3]], ref[:3, [0, 2, 3]])
    np.testing.assert_array_equal(back[:, 1], ref[:3, 1].astype(np.float32))


def test_raw_dat_methods(tmp_path):
    from ct2.tomo3D import image3D, sinogram3D
    im = image3D((3, 4, 5))
    im.mat[:] = np.arange(60, dtype=np.float32).reshape(3, 4, 5)
    p = str(tmp_path / "im.dat")
    im.write_file(p)
    assert open(p, "rb").read() == im.mat.tobytes()
    im2 = image3D((3, 4, 5))
    im2.read(p)
    np.testing.assert_array_equal(im2.mat, im.mat)
    s = sinogram3D(shape=(2, 3, 4))
    s.mat[:] = 7.0
    s.write_file(p)
    s2 = sinogram3D(shape=(2, 3, 4))
    s2.read(p)
    np.testing.assert_array_equal(s2.mat, s.mat)
