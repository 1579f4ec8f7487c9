"""The C2 reconstruction at full size, on the device end to end.

256^3 volume (0.8 mm) of the 40-ellipsoid phantom (seed 20260101) x
0.02/mm, circular cone beam of 360 views, 512x512 detector at 0.8 mm,
R = 500 mm, SDD = 1000 mm, Poisson noise at 1e5 photons per ray (seed
20260102).  FDK (recon3D.py:50-139), the scalar that fits it to the data,
then 20 iterations of EM + 5 normalised TV steps (ct2.solvers.EmTv), the
first eager and the rest replays of one captured HIP graph.

Properties checked (no oracle runs at this size):
* the image is non-negative after every iteration;
* the data residual |A x - b| falls at every iteration;
* relative L2 to the phantom: at most 0.40 after FDK and 0.13 after EM+TV
  (measured 0.358 and 0.117).
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_c2_fdk_em_tv_full_size():
    from ct2 import FDK, phantoms, tomo3D
    from ct2.operator import ConeBeamOperator
    from ct2.solvers import EmTv, capture
    n, ns, nd, vox, pitch, R, SDD, mu = 256, 360, 512, 0.8, 0.8, 500.0, 1000.0, 0.02
    hw = n * vox / 2
    sino = tomo3D.sinogram3D("circular_conebeam", {"radius": R, "source_to_detector": SDD},
                             (ns, nd, nd), 2 * np.pi, nd * pitch, nd * pitch, 0.0,
                             -nd * pitch / 2, -nd * pitch / 2, 0)
    img = tomo3D.image3D((n, n, n), 2 * hw, 2 * hw, 2 * hw, -hw, -hw, -hw)
    op = ConeBeamOperator.from_objects(sino, img, arith="fast")
    truth = torch.from_numpy(phantoms.random_ellipsoids(n, hw, 40, 20260101)).cuda() * mu
    p = op.forward(truth)
    g = torch.Generator(device="cuda").manual_seed(20260102)
    counts = torch.poisson(1e5 * torch.exp(-p.double()), generator=g)
    b = (-torch.log(torch.clamp(counts, min=1.0) / 1e5)).float().contiguous()
    sino.mat = b

    def err(v):
        return float(torch.linalg.vector_norm(v - truth) / torch.linalg.vector_norm(truth))

    x = FDK.fdk_reconstruct(b, sino, img).contiguous()
    x.clamp_(min=0.0)
    e_fdk = err(x)
    ax = op.forward(x)
    x.mul_(float((ax.double() * b.double()).sum() / (ax.double() * ax.double()).sum()))
    st = EmTv(op, b, x, n_tv=5, eps=1e-8)
    resid = [float(torch.linalg.vector_norm(op.forward(x) - b))]
    gr = None
    for k in range(20):
        if gr is None:
            st.iterate()
            gr = capture(st.iterate)
        else:
            gr.replay()
        torch.cuda.synchronize()
        assert float(x.min()) >= 0.0, k
        resid.append(float(torch.linalg.vector_norm(op.forward(x) - b)))
    e_em = err(x)
    print(f"C2: rel L2 to phantom FDK {e_fdk:.4f}, FDK+EM-TV {e_em:.4f}; residual "
          f"{[round(r, 3) for r in resid]}", flush=True)
    assert all(r1 < r0 for r0, r1 in zip(resid, resid[1:])), resid
    assert e_fdk <= 0.40, e_fdk
    assert e_em <= 0.13, e_em
