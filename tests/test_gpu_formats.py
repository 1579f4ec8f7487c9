"""MGHread(device=True): the Fortran-to-C transpose on the GPU gives the same
array as the host reader, for 2-D, 3-D and big-endian files."""
import numpy as np
import pytest

from ct2 import Utils

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,dtype", [((33, 20, 17), np.float32), ((40, 9), np.int16),
                                         ((8, 7, 6), np.float64)])
def test_mgh_device_read(tmp_path, shape, dtype):
    a = (np.random.default_rng(0).random(shape) * 1000).astype(dtype)
    p = str(tmp_path / "a.mgh")
    Utils.MGHwrite(p, a, write_type=dtype)
    host = Utils.MGHread(p)
    dev = Utils.MGHread(p, device=True)
    assert dev.is_cuda and tuple(dev.shape) == shape and dev.is_contiguous()
    np.testing.assert_array_equal(dev.cpu().numpy(), host)
    # big-endian file
    raw = open(p, "rb").read()
    hdr = raw[:512].decode().replace("little_endian", "big_endian   ")
    data = np.frombuffer(raw[512:], dtype=host.dtype.newbyteorder("<")).astype(host.dtype.newbyteorder(">"))
    open(p, "wb").write(hdr.encode() + data.tobytes())
    np.testing.assert_array_equal(Utils.MGHread(p, device=True).cpu().numpy(), host)
