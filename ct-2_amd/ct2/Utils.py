"""On-disk formats of ct-2 (emil_code/Utils.py and the example drivers).

* MGH volumes (Utils.py:7-135): a 512-byte ASCII header of ``KEY= value;``
  lines inside ``{ }``, space padded, then the array in Fortran order.
  ``MGHread`` returns the C-ordered array (or a device tensor with
  ``device=True``; the Fortran-to-C transpose then runs on the GPU).
  ``MGHwrite``, ``WriteSino`` and ``WriteImage`` write the reference's
  header lines in insertion order, then the data.  Unlike the reference, the reader
  honours ``BYTE_ORDER= big_endian`` and the writer refuses a header that
  does not fit in 512 bytes (the reference would silently shift the data).
* ``GEdata2sino`` (Utils.py:177-237): a GE tomosynthesis projection stack
  as a ``sinogram3D`` with the scanner's 11 angles.
* Raw ``.dat`` arrays (tomo3D.py:202-208, :254-257, :876-885): native
  float32 bytes in C order -- ``write_dat`` / ``read_dat``.
* Run state of the iterative drivers (examples/jawPhantomCCB.py:148-168,
  examples/sphericalRadon.py:171-183): ``imageTemp.dat``, ``iternum.txt``
  (the iteration number) and ``parms.txt`` (gf, reg_fac, meps, one per
  line).  ``read_state`` reads parms.txt line by line; the reference calls
  ``read()`` three times on one handle, so its second and third reads come
  back empty and the resume fails (jawPhantomCCB.py:160-162).
* ``derrEvol.txt`` (jawPhantomCCB.py:257-261): one tab-separated line per
  iteration, ``iter  data-distance  TV  image-error``.
"""
from __future__ import annotations

import os

import numpy as np

HEADER_BYTES = 512
_TYPES = {"unsigned_short": np.uint16, "short": np.int16, "float": np.float32,
          "double": np.float64}
_NAMES = {np.dtype(np.float32): "float", np.dtype(np.float64): "double",
          np.dtype(np.uint16): "unsigned_short", np.dtype(np.int16): "short"}


# -------------------------------------------------------------------- MGH ----
def MGHheader(filename):
    """The header items of an MGH file as a dict (numeric values as int)."""
    with open(filename, "rb") as f:
        text = f.read(HEADER_BYTES).decode("latin-1")
    items = {}
    for line in text.splitlines():
        line = line.strip()
        if not line or line in "{}" or not line.endswith(";"):
            continue
        key, _, val = line[:-1].partition(" ")
        val = val.strip()
        items[key] = int(val) if val.isdigit() else val
    return items


def MGHread(filename, device=False):
    """Utils.py:7-66: the array of an MGH file in C order (shape SIZE1, SIZE2
    [, SIZE3]).  With device=True, a torch tensor on the current HIP device."""
    h = MGHheader(filename)
    t = h.get("TYPE=")
    if t not in _TYPES:
        raise ValueError(f"{filename}: unsupported MGH TYPE= {t!r}")
    dt = np.dtype(_TYPES[t])
    if str(h.get("BYTE_ORDER=", "little_endian")).startswith("big"):
        dt = dt.newbyteorder(">")
    dim = int(h.get("DIM=", 1))
    sizes = [int(h[f"SIZE{i}="]) for i in range(1, dim + 1)] if dim in (2, 3) else None
    a = np.fromfile(filename, dtype=dt, offset=int(h.get("HEADER_BYTES=", HEADER_BYTES)))
    if sizes is None:
        return a.astype(dt.newbyteorder("="))
    n = int(np.prod(sizes))
    if a.size < n:
        raise ValueError(f"{filename}: {a.size} elements for sizes {sizes}")
    a = a[:n].reshape(sizes[::-1])            # Fortran layout read as C
    if device:
        import torch
        from . import _dev
        host = np.ascontiguousarray(a.astype(dt.newbyteorder("=")))
        t = torch.from_numpy(host).pin_memory().to(_dev.device(), non_blocking=True)
        return t.permute(*range(dim - 1, -1, -1)).contiguous()
    return np.ascontiguousarray(a.transpose()).astype(dt.newbyteorder("="), copy=False)


def _header_text(items):
    s = "{\n" + "".join(f"{k} {v};\n" for k, v in items.items()) + "}\n\n"
    if len(s) > HEADER_BYTES:
        raise ValueError(f"MGH header of {len(s)} bytes does not fit in {HEADER_BYTES}")
    return s + " " * (HEADER_BYTES - len(s))


def MGHwrite(filename, a, d=None, write_type=''):
    """Utils.py:69-130: float32, float64, int16 or uint16 data; int64 is
    written as int16 and float64 as float32, as in the reference."""
    if not isinstance(a, np.ndarray):
        a = a.detach().cpu().numpy() if hasattr(a, "detach") else np.asarray(a)
    if write_type != '':
        a = a.astype(write_type)
    if a.dtype == np.int64:
        a = a.astype(np.int16)
    if a.dtype == np.float64:
        a = a.astype(np.float32)
    if a.dtype not in _NAMES:
        raise ValueError(f"MGHwrite: unsupported dtype {a.dtype}")
    items = {"HEADER_BYTES=": HEADER_BYTES, "DIM=": a.ndim, "SIZE1=": a.shape[0],
             "SIZE2=": a.shape[1], "BYTE_ORDER=": "little_endian", "TYPE=": _NAMES[a.dtype]}
    if a.ndim == 3:
        items["SIZE3="] = a.shape[2]
    items.update(d or {})
    with open(filename, "wb") as f:
        f.write(_header_text(items).encode("latin-1"))
        a.astype(a.dtype.newbyteorder("<")).transpose().tofile(f)


def _host(m):
    return m.detach().cpu().numpy() if hasattr(m, "detach") else np.asarray(m)


def WriteSino(filename, sino, comment_line=''):
    """Utils.py:133-152: the sinogram with its geometry in the header."""
    parms = dict(img_type='sinogram3D', config_name=sino.config_name,
                 shape=tuple(_host(sino.mat).shape), slen=sino.slen, s0=sino.s0,
                 ulen=sino.ulen, u0=sino.u0, vlen=sino.vlen, v0=sino.v0,
                 s_include_endpoints=sino.s_include_endpoints, comment=comment_line)
    parms.update(sino.parms)
    MGHwrite(filename, _host(sino.mat), parms)


def WriteImage(filename, image, comment_line=''):
    """Utils.py:154-173."""
    parms = dict(img_type='image3D', shape=tuple(_host(image.mat).shape), xlen=image.xlen,
                 ylen=image.ylen, zlen=image.zlen, x0=image.x0, y0=image.y0, z0=image.z0,
                 comment=comment_line)
    MGHwrite(filename, _host(image.mat), parms)


# GE tomosynthesis acquisition angles in degrees (Utils.py:195, :205)
GE_ANGLES_R = (-24.98, -19.52, -14.19, -8.99, -3.90, 1.09, 5.99, 10.82, 15.59, 20.32, 25.00)
GE_ANGLES_L = (-25.00, -20.32, -15.59, -10.82, -5.99, -1.09, 3.90, 8.99, 14.19, 19.52, 24.98)


def GEdata2sino(filename, pixsize=0.01, Laterality='n'):
    """Utils.py:177-237: GE projection stack -> tomosynthesis sinogram3D
    (centre-to-detector 23.7, source radius 44.3, negative counts zeroed).
    Laterality 'R' keeps the detector rows, 'L' reverses them."""
    from . import tomo3D as t3
    img = MGHread(filename).transpose(2, 0, 1)
    if Laterality == 'R':
        img = np.ascontiguousarray(img, dtype=np.float32)
        ang = np.array(GE_ANGLES_R)
    elif Laterality == 'L':
        img = np.ascontiguousarray(img[:, ::-1, :], dtype=np.float32)
        ang = np.array(GE_ANGLES_L)
    else:
        raise ValueError("Laterality must be 'R' or 'L'")
    ang = (ang + 90) / 180. * np.pi
    if img.shape[0] != ang.size:
        raise ValueError(f"{filename}: {img.shape[0]} projections, expected {ang.size}")
    ulen0, vlen0 = img.shape[1] * pixsize, img.shape[2] * pixsize
    sino = t3.sinogram3D(config_name='tomosynthesis',
                         parms={"center_to_detector": 23.7, "radius": 44.3},
                         shape=img.shape, slen=img.shape[0], s0=np.pi / 4.,
                         ulen=ulen0, u0=0., vlen=vlen0, v0=-vlen0 / 2.)
    img[img < 0.0] = 0.0
    sino.mat = img
    fv = np.array(sino.frame_vectors, np.float64, copy=True)
    for ip in range(img.shape[0]):
        fv[ip] = sino.frame(ang[ip], sino.parms)
    sino.frame_vectors = fv.astype(np.float32)
    return sino


# ------------------------------------------------------ raw arrays, state ----
def write_dat(filename, mat):
    """tomo3D.py:254-257: raw native bytes in C order."""
    _host(mat).tofile(filename, sep='', format='')


def read_dat(filename, shape, dtype=np.float32):
    """tomo3D.py:202-208: the inverse of write_dat."""
    a = np.fromfile(filename, dtype=dtype)
    return a.reshape(shape)


def write_state(base_dir, it, gf, reg_fac, meps, mat=None):
    """examples/sphericalRadon.py:171-183: imageTemp.dat, iternum.txt and
    parms.txt, each replaced atomically."""
    os.makedirs(base_dir, exist_ok=True)

    def put(name, data, mode="w"):
        tmp = os.path.join(base_dir, name + ".tmp")
        if mode == "w":
            with open(tmp, "w") as f:
                f.write(data)
        else:
            write_dat(tmp, data)
        os.replace(tmp, os.path.join(base_dir, name))

    if mat is not None:
        put("imageTemp.dat", mat, "b")
    put("iternum.txt", str(int(it)))
    put("parms.txt", f"{float(gf)!r}\n{float(reg_fac)!r}\n{float(meps)!r}\n")


def read_state(base_dir):
    """(iter, gf, reg_fac, meps) from iternum.txt and parms.txt."""
    with open(os.path.join(base_dir, "iternum.txt")) as f:
        it = int(f.read().strip())
    with open(os.path.join(base_dir, "parms.txt")) as f:
        vals = [float(x) for x in f.read().split()]
    if len(vals) != 3:
        raise ValueError(f"{base_dir}/parms.txt: expected gf, reg_fac, meps; got {vals}")
    return it, vals[0], vals[1], vals[2]


def append_derr(base_dir, it, ddist, tv, ierr):
    """One line of derrEvol.txt (jawPhantomCCB.py:257-261)."""
    os.makedirs(base_dir, exist_ok=True)
    with open(os.path.join(base_dir, "derrEvol.txt"), "a") as f:
        f.write(f"{int(it)}\t{float(ddist)!r}\t{float(tv)!r}\t{float(ierr)!r}\n")


def read_derr(filename):
    """derrEvol.txt as an (n, 4) float64 array: iter, data distance, TV,
    image error."""
    rows = [ln.split() for ln in open(filename) if ln.strip()]
    return np.array([[float(x) for x in r[:4]] for r in rows], np.float64).reshape(-1, 4)
