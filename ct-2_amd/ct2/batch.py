"""Batched plans: many independent cone-beam problems of uneven size (volume
shape, view count, detector) projected, backprojected or reconstructed by
SART in single persistent launches (ct2_batch_* in libct2.so).

The reference handles one problem per call (emil_code/tomo3D.py:1127-1168,
:432-472); a plan returns, per problem, what ConeBeamOperator.forward /
adjoint / sart_step return, with every problem advanced by one launch.  SART
walks ``nsub`` ordered subsets of interleaved views (subset s = views s,
s+nsub, ...).  After creation no call allocates or synchronises, so a loop of
``sart`` calls can be captured into a HIP graph (ct2.solvers.capture).
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _dev
from ._lib import call, ct2_desc, lib


class BatchPlan:
    def __init__(self, ops, sinos, vols, nsub=10, proj=None, adj=None, tune=True):
        """ops: ConeBeamOperator per problem (geometry, frame vectors,
        indicator; all with the same interpolation and arithmetic);
        sinos / vols: contiguous float32 device tensors of each problem's
        sinogram and volume; proj / adj: optional output tensors of
        ``project`` and ``backproject``."""
        n = len(ops)
        if not (len(sinos) == len(vols) == n) or n == 0:
            raise ValueError("one sinogram and one volume per operator")
        for i, (op, s, v) in enumerate(zip(ops, sinos, vols)):
            if tuple(s.shape) != op.sino_shape or tuple(v.shape) != op.shape:
                raise ValueError(f"problem {i}: array shapes do not match the operator")
            for t in (s, v):
                if t.dtype != torch.float32 or not t.is_contiguous() or not t.is_cuda:
                    raise ValueError(f"problem {i}: arrays must be contiguous float32 device tensors")
        self.ops = list(ops)
        # the library keeps raw pointers: hold every tensor for the plan's life
        self._keep = [list(sinos), list(vols), list(proj or []), list(adj or []),
                      [op.fv for op in ops], [op.ind for op in ops]]
        descs = (ct2_desc * n)()
        for i, op in enumerate(ops):
            descs[i] = op._desc(torch.float32)

        def ptrs(ts):
            if ts is None:
                return None
            return (C.c_void_p * n)(*[_dev.ptr(t).value if t is not None else None for t in ts])

        self._h = C.c_void_p()
        call("ct2_batch_create", C.byref(self._h), n, descs, ptrs([op.fv for op in ops]),
             ptrs([op.ind for op in ops]), ptrs(sinos), ptrs(vols), ptrs(proj), ptrs(adj),
             int(nsub), int(bool(tune)), _dev.stream_handle())
        self.nsub = int(nsub)

    def info(self):
        """Tuned shapes and memory: dict(tile, fwd_grid, brick, bp_grid, nsub, nprob, bytes)."""
        info = (C.c_int * 8)()
        nbytes = C.c_size_t()
        call("ct2_batch_info", self._h, info, C.byref(nbytes))
        return {"tile": (info[0], 64), "fwd_grid": info[1], "brick": (info[2], info[3], info[4]),
                "bp_grid": info[5], "nsub": info[6], "nprob": info[7], "bytes": nbytes.value}

    def project(self):
        call("ct2_batch_project", self._h, _dev.stream_handle())

    def backproject(self):
        call("ct2_batch_backproject", self._h, _dev.stream_handle())

    def sart(self, lam=1.0, niter=1):
        call("ct2_batch_sart", self._h, float(lam), int(niter), _dev.stream_handle())

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().ct2_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def updates_per_sart_iteration(self):
        """Voxel-ray updates of one SART iteration: a forward and an adjoint
        over every view of every problem."""
        return 2 * sum(op.updates_per_direction() for op in self.ops)
