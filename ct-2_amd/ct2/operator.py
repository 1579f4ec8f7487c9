"""Device-resident Joseph operator: the forward projector A and its exact
adjoint A^T for one scan geometry, plus the update steps built on them.

Everything stays in HBM as torch tensors; calls are asynchronous on the
current stream and allocate only their outputs (or nothing, given ``out=``).
Arithmetic ``"exact"`` reproduces the reference's FP32 operation order
(emil_code/view_proj3D.f:703-910, bit-identical forward); ``"fast"`` sets
each ray up in FP64 and steps planes with FP32 FMAs and compensated sums.
Volume storage may be float32, float16 or bfloat16 for the forward
projection; the adjoint accumulates in FP64 in LDS and writes the volume in
the requested storage type.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _dev
from ._lib import (CT2_ARITH_EXACT, CT2_ARITH_FAST, CT2_AXIS_FREE, CT2_AXIS_Z, CT2_BF16,
                   CT2_BF16_ZPLANES, CT2_F16, CT2_F16_ZPLANES, CT2_F16Z2, CT2_F32,
                   CT2_F32_ZPLANES, CT2_LINEAR, CT2_NEAREST, CT2_ROUND_NEAREST,
                   CT2_ROUND_STOCHASTIC, call, ct2_desc, lib)

_DT = {torch.float32: CT2_F32, torch.float16: CT2_F16, torch.bfloat16: CT2_BF16}
_DTZ = {torch.float32: CT2_F32_ZPLANES, torch.float16: CT2_F16_ZPLANES,
        torch.bfloat16: CT2_BF16_ZPLANES}


def to_storage(vol32, dtype=torch.float16, rounding="stochastic", seed=20260107, out=None,
               pairs=False, zplanes=False):
    """float32 volume -> 16-bit storage (float16 or bfloat16) by the library's
    conversion kernel; ``rounding`` is "stochastic" (unbiased, deterministic
    in ``seed``) or "nearest".  ``pairs`` (float16 only) gives the f16 pair
    storage read by the fast forward projector: a (*shape, 2) tensor whose
    [..., 0] is the volume as float16 storage rounds it and [..., 1] the next
    element along the last axis (zero after the last element).  ``zplanes``
    gives the plane-major copy (dtype float32, float16 or bfloat16): an (nz,
    nx, ny) tensor, element [k, i, j] = voxel (i, j, k) rounded as the C-order
    conversion rounds it, for ``ConeBeamOperator.forward(v, layout="zxy")``."""
    if vol32.dtype != torch.float32 or not vol32.is_contiguous():
        raise ValueError("source must be a contiguous float32 tensor")
    if zplanes:
        if vol32.dim() != 3 or pairs:
            raise ValueError("plane-major storage takes a 3-D volume (no pairs)")
        nx, ny, nz = vol32.shape
        if dtype not in _DTZ:
            raise ValueError("plane-major storage dtype must be float32, float16 or bfloat16")
        shape = (nz, nx, ny)
        out = torch.empty(shape, dtype=dtype, device=vol32.device) if out is None else out
        if tuple(out.shape) != shape or out.dtype != dtype or not out.is_contiguous():
            raise ValueError(f"output must be a contiguous {dtype} tensor of shape {shape}")
        call("ct2_convert_volume_zplanes", _dev.ptr(vol32), _dev.ptr(out), _DTZ[dtype], nx, ny, nz,
             {"stochastic": CT2_ROUND_STOCHASTIC, "nearest": CT2_ROUND_NEAREST}[rounding],
             int(seed), _dev.stream_handle())
        return out
    if dtype not in (torch.float16, torch.bfloat16):
        raise ValueError("storage dtype must be float16 or bfloat16")
    if pairs and dtype != torch.float16:
        raise ValueError("pair storage is float16")
    shape = tuple(vol32.shape) + ((2,) if pairs else ())
    out = torch.empty(shape, dtype=dtype, device=vol32.device) if out is None else out
    if tuple(out.shape) != shape or out.dtype != dtype or not out.is_contiguous():
        raise ValueError(f"output must be a contiguous {dtype} tensor of shape {shape}")
    call("ct2_convert_volume", _dev.ptr(vol32), _dev.ptr(out), CT2_F16Z2 if pairs else _DT[dtype],
         C.c_size_t(vol32.numel()),
         {"stochastic": CT2_ROUND_STOCHASTIC, "nearest": CT2_ROUND_NEAREST}[rounding],
         int(seed), _dev.stream_handle())
    return out


class ConeBeamOperator:
    def __init__(self, frame_vectors, ns, nu, nv, du, dv, u0, v0, shape, voxel, origin,
                 interp="linear", axis="free", arith="fast", indicator=None, brick=None):
        nx, ny, nz = shape
        dx, dy, dz = voxel
        x0, y0, z0 = origin
        self.shape = (nx, ny, nz)
        self.sino_shape = (ns, nu, nv)
        self.d = ct2_desc()
        call("ct2_desc_init", C.byref(self.d), ns, nu, nv, du, dv, u0, v0, nx, ny, nz,
             dx, dy, dz, x0, y0, z0)
        self.d.interp = {"linear": CT2_LINEAR, "nearest": CT2_NEAREST}[interp]
        self.d.axis = {"free": CT2_AXIS_FREE, "z": CT2_AXIS_Z}[axis]
        self.d.arith = {"exact": CT2_ARITH_EXACT, "fast": CT2_ARITH_FAST}[arith]
        if brick is not None:
            self.d.brick[0], self.d.brick[1], self.d.brick[2] = brick
        self.fv = _dev.to_dev(frame_vectors, torch.float32, (ns, 18))
        self.ind = _dev.ind_dev(indicator, (ns, nu, nv))
        self.device = self.fv.device
        self._sens = None
        self._work = None   # the column-separable adjoint's tables (bytes)

    @classmethod
    def from_objects(cls, sino, image, **kw):
        """Operator for a tomo3D sinogram3D / image3D pair."""
        return cls(sino.frame_vectors, sino.ns, sino.nu, sino.nv, sino.du, sino.dv, sino.u0,
                   sino.v0, (image.nx, image.ny, image.nz), (image.dx, image.dy, image.dz),
                   (image.x0, image.y0, image.z0), indicator=sino.indicator, **kw)

    def _desc(self, dtype=torch.float32, view0=0, nviews=0, vstride=1):
        d = ct2_desc.from_buffer_copy(self.d)
        d.vol_dtype = dtype if isinstance(dtype, int) else _DT[dtype]
        d.view0 = view0
        d.nviews = nviews
        d.vstride = vstride
        return d

    def forward(self, vol, out=None, view0=0, nviews=0, vstride=1, layout="xyz"):
        """A vol -> sinogram (float32).  vol: (nx,ny,nz) float32/float16/bfloat16,
        or (nx,ny,nz,2) float16 pair storage (``to_storage(..., pairs=True)``),
        which projects exactly as its float16 values do; with layout="zxy" the
        plane-major (nz,nx,ny) storage of ``to_storage(..., zplanes=True)``,
        which projects exactly as the same values in (nx,ny,nz) order do.
        With a view subset (views view0 + j*vstride, j < nviews; nviews 0 = all
        such views) only those views of ``out`` are written."""
        nx, ny, nz = self.shape
        if layout == "zxy":
            if tuple(vol.shape) != (nz, nx, ny) or vol.dtype not in _DTZ:
                raise ValueError(f"plane-major volume must be a {(nz, nx, ny)} "
                                 "float32/16/bf16 tensor")
            code = _DTZ[vol.dtype]
        elif layout == "xyz":
            pairs = tuple(vol.shape) == self.shape + (2,) and vol.dtype == torch.float16
            if not pairs and (vol.shape != self.shape or vol.dtype not in _DT):
                raise ValueError(f"volume must be a {self.shape} float32/16/bf16 tensor or its "
                                 f"{self.shape + (2,)} float16 pair storage")
            code = CT2_F16Z2 if pairs else vol.dtype
        else:
            raise ValueError('layout must be "xyz" or "zxy"')
        if not vol.is_contiguous():
            raise ValueError("volume must be contiguous")
        if out is None:
            out = torch.empty(self.sino_shape, dtype=torch.float32, device=vol.device)
        d = self._desc(code, view0, nviews, vstride)
        call("ct2_project", C.byref(d), _dev.ptr(out), _dev.ptr(self.ind), _dev.ptr(self.fv),
             _dev.ptr(vol), _dev.stream_handle())
        return out

    def adjoint(self, sino, out=None, dtype=torch.float32, view0=0, nviews=0, vstride=1):
        """A^T sino -> volume written in `dtype` storage (over a view subset as
        in ``forward``).  The column-separable route keeps its column tables
        in a workspace the operator allocates once and reuses, so calls on
        one operator are ordered on one stream (or serialised by the
        caller)."""
        if sino.shape != self.sino_shape or sino.dtype != torch.float32 or not sino.is_contiguous():
            raise ValueError(f"sinogram must be a contiguous float32 {self.sino_shape} tensor")
        if out is None:
            out = torch.empty(self.shape, dtype=dtype, device=sino.device)
        d = self._desc(out.dtype, view0, nviews, vstride)
        n = int(lib().ct2_backproject_work_bytes(C.byref(d)))
        if n and (self._work is None or self._work.numel() < n):
            self._work = torch.empty(n, dtype=torch.uint8, device=sino.device)
        call("ct2_backproject_ws", C.byref(d), _dev.ptr(out), _dev.ptr(sino), _dev.ptr(self.ind),
             _dev.ptr(self.fv), _dev.ptr(self._work) if n else None, C.c_size_t(n),
             _dev.stream_handle())
        return out

    def sensitivity(self):
        """A^T 1 over the used rays (cached)."""
        if self._sens is None:
            ones = torch.ones(self.sino_shape, dtype=torch.float32, device=self.device)
            self._sens = self.adjoint(ones)
        return self._sens

    def em_step(self, sino, vol, work=None, use_cached_sensitivity=True):
        """MLEM update of vol in place (view_proj3D.f:2507-2831 semantics)."""
        n = int(np.prod(self.sino_shape)) * 2 + int(np.prod(self.shape)) * 2
        if work is None:
            work = torch.empty(n, dtype=torch.float32, device=vol.device)
        sens = self.sensitivity() if use_cached_sensitivity else None
        d = self._desc()
        call("ct2_em_step", C.byref(d), _dev.ptr(sino), _dev.ptr(self.ind), _dev.ptr(self.fv),
             _dev.ptr(vol), _dev.ptr(work), _dev.ptr(sens), _dev.stream_handle())
        return vol

    def sart_step(self, sino, vol, lam=1.0, view0=0, nviews=0, work=None, vstride=1):
        """SART update of vol in place over the views view0 + j*vstride."""
        n = int(np.prod(self.sino_shape)) * 2 + int(np.prod(self.shape)) * 3
        if work is None:
            work = torch.empty(n, dtype=torch.float32, device=vol.device)
        d = self._desc(torch.float32, view0, nviews, vstride)
        call("ct2_sart_step", C.byref(d), float(lam), _dev.ptr(sino), _dev.ptr(self.ind),
             _dev.ptr(self.fv), _dev.ptr(vol), _dev.ptr(work), _dev.stream_handle())
        return vol

    def art_step(self, sino, vol, reg_fac=1.0):
        """Sequential Kaczmarz sweep, order-exact (view_proj3D.f:1950-2271)."""
        d = self._desc()
        call("ct2_art_step", C.byref(d), float(reg_fac), _dev.ptr(sino), _dev.ptr(self.ind),
             _dev.ptr(self.fv), _dev.ptr(vol), _dev.stream_handle())
        return vol

    def updates_per_direction(self):
        """Voxel-ray updates of one projection: sum over rays of the driving-axis
        length (ns*nu*nv*n for cubic volumes)."""
        ns, nu, nv = self.sino_shape
        return ns * nu * nv * max(self.shape)
