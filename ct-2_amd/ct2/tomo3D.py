"""Object model of the reference (emil_code/tomo3D.py), Python 3: the call
sites of the reconstruction hot path.

``sinogram3D``, ``image3D`` and ``linogram3D`` keep the reference's
constructors, attributes (``mat``, ``indicator``, ``frame_vectors``,
``ray_vectors``, ``ds/du/dv``, ``dx/dy/dz``...) and methods, and dispatch
through the same algorithm-name dictionaries (emil_code/tomo3D.py:55-103) to
the drop-in modules of this package.  ``mat`` may be a numpy array (each call
round-trips to the device, as the reference round-trips through f2py) or a
torch device tensor (everything stays in HBM); use ``to_device()``.
Bad algorithm names raise ValueError (the reference calls ``sys.exit`` without
importing ``sys``, emil_code/tomo3D.py:15, :440-441).
"""
from __future__ import annotations

import numpy as np

from . import FDK as _FDK  # noqa: F401  (the FDK call surface lives next to the model)
from . import configs3D, ray_proj3D, totalvar, view_proj3D
from . import view_proj3D_tomo as vpt
from .quadric import (conical, cylindrical, ellipsoid, ellipsoidal, gen_object,  # noqa: F401
                      phantom3D, planar, shape3D, surface)

pi = np.pi
default_type = np.float32
indicator_type = np.int8
default_data_shape = 32, 128, 128
default_image_shape = 128, 128, 128
default_num_rays = 1000000

bp_algs_dict = {
    "adj_nearest": view_proj3D.sinobackproject_nearest,
    "adj_linear": view_proj3D.sinobackproject_linear,
    "pd_nearest": view_proj3D.pd_backproj_nearest,
    "pd_linear": view_proj3D.pd_backproj_linear,
}
proj_algs_dict = {
    "proj_linear": view_proj3D.sinoproj_linear,
    "proj_nearest": view_proj3D.sinoproj_nearest,
    "zproj_linear": vpt.sinoprojz,
}
art_algs_dict = {
    "art_linear": view_proj3D.artstep_linear,
    "art_nearest": view_proj3D.artstep_nearest,
    "zart_linear": vpt.artstepz,
}
em_algs_dict = {
    "em_linear": view_proj3D.emstep_linear,
    "em_nearest": view_proj3D.emstep_nearest,
    "zem_linear": vpt.emstep_linearz,
}
ray_bp_algs_dict = {"adj_nearest": ray_proj3D.sinobackproject_nearest}
ray_proj_algs_dict = {"proj_nearest": ray_proj3D.sinoproj_nearest,
                      "proj_linear": ray_proj3D.sinoproj_linear}
ray_art_algs_dict = {"art_nearest": ray_proj3D.artstep_nearest,
                     "art_linear": ray_proj3D.artstep_linear}
ray_em_algs_dict = {"em_nearest": ray_proj3D.emstep_nearest}


def _round_half_away(x):
    """Python 2's round(): halves go away from zero (Python 3 rounds them to
    even); returns a float."""
    return float(np.copysign(np.floor(abs(x) + 0.5), x))


def _pick(table, alg, what):
    if alg not in table:
        raise ValueError(f"bad algorithm choice for {what}: {alg!r} (choose from {sorted(table)})")
    return table[alg]


def _torch():
    import torch
    return torch


def _is_tensor(x):
    return type(x).__module__.startswith("torch")


def _xp_sum_sq(a, b, w=None):
    """sqrt(sum(w*(a-b)^2)) for numpy or torch, as emil_code/tomo3D.py:319-324
    and :946-951 evaluate it: the difference and its square in the arrays'
    own precision (float32 for float32 images and sinograms), the weight
    (0 or 1, so exact) and the sum in float64."""
    if _is_tensor(a) or _is_tensor(b):
        torch = _torch()
        a = torch.as_tensor(a)
        b = torch.as_tensor(b, device=a.device)
        d = a - b
        d = (d * d).to(torch.float64)
        if w is not None:
            d = d * torch.as_tensor(w, device=a.device).to(torch.float64)
        return float(torch.sqrt(d.sum()))
    d = np.asarray(a).ravel() - np.asarray(b).ravel()
    d = (d * d).astype(np.float64)
    if w is not None:
        d = d * np.asarray(w, np.float64).ravel()
    return float(np.sqrt(d.sum(dtype=np.float64)))


class sinogram3D:
    """3D sinogram array (emil_code/tomo3D.py:110-476)."""

    def __init__(self, config_name='circular_conebeam',
                 parms={"source_to_detector": 10., "radius": 5.},
                 shape=default_data_shape, slen=pi, ulen=4.0, vlen=4.0,
                 s0=0., u0=-2.0, v0=-2.0, s_include_endpoints=1):
        self.ns, self.nu, self.nv = shape
        self.mat = np.zeros(shape, default_type)
        self.indicator = np.ones(shape, indicator_type)
        self.config_name = config_name
        self.parms = parms
        self.frame = configs3D.configs[config_name]
        self.slen, self.ulen, self.vlen = slen, ulen, vlen
        self.s_include_endpoints = s_include_endpoints
        self.s0, self.u0, self.v0 = s0, u0, v0
        self.ds = self.get_ds()
        self.du = self.get_du()
        self.dv = self.get_dv()
        self.frame_vectors = configs3D.frame_vectors(config_name, parms, self.ns, slen, s0,
                                                     s_include_endpoints)

    def get_ds(self):
        return self.slen / (self.ns - 1.) if self.s_include_endpoints else self.slen / self.ns

    def get_du(self):
        return self.ulen / self.nu

    def get_dv(self):
        return self.vlen / self.nv

    def get_parms(self):
        return self.parms

    def frame_vecs(self, s):
        c = self.frame(s, self.parms)
        return (np.array(c[0:3]), np.array(c[3:6]), np.array(c[6:9]), np.array(c[9:12]),
                np.array(c[12:15]), np.array(c[15:18]))

    def read(self, file_name, data_type=default_type):
        """tomo3D.py:202-208 / :876-880: raw array of the current shape; a
        device array stays on the device."""
        shape = tuple(self.mat.shape)
        a = np.fromfile(file_name, dtype=data_type).reshape(shape)
        if _is_tensor(self.mat):
            self.mat = _torch().from_numpy(a).to(self.mat.device)
        else:
            self.mat = a

    def write_file(self, file_name):
        """tomo3D.py:254-257 / :883-885: raw native bytes in C order."""
        from . import Utils
        Utils.write_dat(file_name, self.mat)

    def duplicate(self):
        new = sinogram3D(self.config_name, self.parms, (self.ns, self.nu, self.nv), self.slen,
                         self.ulen, self.vlen, self.s0, self.u0, self.v0, self.s_include_endpoints)
        new.mat = self.mat.clone() if _is_tensor(self.mat) else self.mat.copy()
        new.indicator = self.indicator.clone() if _is_tensor(self.indicator) else self.indicator.copy()
        new.frame_vectors = self.frame_vectors
        return new

    def to_device(self):
        """Move mat, indicator and frame vectors to the current HIP device."""
        torch = _torch()
        from . import _dev
        self.mat = _dev.to_dev(self.mat, torch.float32, (self.ns, self.nu, self.nv))
        self.indicator = _dev.ind_dev(self.indicator, (self.ns, self.nu, self.nv))
        self.frame_vectors = _dev.to_dev(self.frame_vectors, torch.float32, (self.ns, 18))
        return self

    def clear(self):
        self.mat *= 0.

    def dist_to(self, sino):
        """emil_code/tomo3D.py:319-324 (float64 norm over used rays)."""
        return _xp_sum_sq(self.mat, sino.mat, self.indicator)

    def copy_to(self, dest):
        dest.mat = self.mat.clone() if _is_tensor(self.mat) else self.mat.copy()
        dest.indicator = self.indicator.clone() if _is_tensor(self.indicator) else self.indicator.copy()

    def flip_detector(self):
        """Exchange the detector's u and v axes (tomo3D.py:213-238): e_u and e_v
        of every frame swap, mat and indicator are transposed to [ns][nv][nu],
        and u0/v0, ulen/vlen, du/dv, nu/nv swap.  The rays are the same, so
        projections come out transposed and backprojections unchanged."""
        fv = self.frame_vectors
        if _is_tensor(fv):
            fv = fv.clone()
            fv[:, 9:12], fv[:, 12:15] = self.frame_vectors[:, 12:15], self.frame_vectors[:, 9:12]
        else:
            fv = np.array(fv)
            fv[:, 9:12], fv[:, 12:15] = self.frame_vectors[:, 12:15], self.frame_vectors[:, 9:12]
        self.frame_vectors = fv
        if _is_tensor(self.mat):
            self.mat = self.mat.transpose(1, 2).contiguous()
        else:
            self.mat = np.ascontiguousarray(self.mat.transpose((0, 2, 1)))
        if _is_tensor(self.indicator):
            self.indicator = self.indicator.transpose(1, 2).contiguous()
        else:
            self.indicator = np.ascontiguousarray(self.indicator.transpose((0, 2, 1)))
        self.shape = tuple(self.mat.shape)
        self.nu, self.nv = self.shape[1], self.shape[2]
        self.u0, self.v0 = self.v0, self.u0
        self.ulen, self.vlen = self.vlen, self.ulen
        self.du, self.dv = self.dv, self.du

    def scramble(self, spacing):
        """Reorder the views' frame vectors as views p, p+spacing, ... for
        p = 0..spacing-1 (tomo3D.py:240-252); ns must be a multiple of
        spacing."""
        if self.ns % spacing:
            raise ValueError(f"ns = {self.ns} is not a multiple of spacing {spacing}")
        idx = np.arange(self.ns).reshape(self.ns // spacing, spacing).T.ravel()
        if _is_tensor(self.frame_vectors):
            idx = _torch().from_numpy(idx).to(self.frame_vectors.device)
        self.frame_vectors = self.frame_vectors[idx]

    def clear_indicator(self):
        """Every ray off (tomo3D.py:262-263)."""
        if _is_tensor(self.indicator):
            self.indicator = self.indicator.new_zeros((self.ns, self.nu, self.nv))
        else:
            self.indicator = np.zeros((self.ns, self.nu, self.nv), indicator_type)

    def config_indicator(self, slow, shigh, ulow, uhigh, vlow, vhigh, ray_op='add'):
        """Turn on the rays in the s-, u- and v-ranges (tomo3D.py:267-299), with
        the reference's grid snapping and loops: Python 2 rounding (half away
        from zero), float accumulation of s, u and v, and v's range open at
        the top while s and u include theirs."""
        if ray_op == 'replace':
            self.clear_indicator()
        r = _round_half_away
        sl = max(self.s0, self.s0 + self.ds * r((slow - self.s0) / self.ds))
        sh = min(self.s0 + self.ds * (self.ns - 1), self.s0 + self.ds * r((shigh - self.s0) / self.ds))
        ul = max(self.u0, self.u0 + self.du * r((ulow - self.u0) / self.du))
        uh = min(self.u0 + self.du * (self.nu - 1), self.u0 + self.du * r((uhigh - self.u0) / self.du))
        vl = max(self.v0, self.v0 + self.dv * r((vlow - self.v0) / self.dv))
        vh = min(self.v0 + self.dv * (self.nv - 1), self.v0 + self.dv * r((vhigh - self.v0) / self.dv))
        on = np.zeros((self.ns, self.nu, self.nv), bool)
        s = sl
        while s <= sh:
            i = int(r((s - self.s0) / self.ds))
            u = ul
            while u <= uh:
                j = int(r((u - self.u0) / self.du))
                v = vl
                while v < vh:
                    on[i, j, int(r((v - self.v0) / self.dv))] = True
                    v += self.dv
                u += self.du
            s += self.ds
        if _is_tensor(self.indicator):
            self.indicator[_torch().from_numpy(on).to(self.indicator.device)] = 1
        else:
            self.indicator[on] = 1

    def _host_mat(self):
        m = self.mat.cpu().numpy() if _is_tensor(self.mat) else self.mat
        return np.asarray(m)

    def _set_mat(self, m):
        m = np.asarray(m, default_type)
        self.mat = _torch().from_numpy(m).to(self.mat.device) if _is_tensor(self.mat) else m

    def _used(self):
        ind = self.indicator.cpu().numpy() if _is_tensor(self.indicator) else self.indicator
        return np.asarray(ind) != 0

    def add_noise(self, std=0.1, rng=None):
        """Multiplicative noise: every bin times max(0, N(1, std))
        (tomo3D.py:334-342).  Draws come in the reference's (i, j, k) loop
        order from ``rng`` (default: numpy's global generator, as the
        reference), so a seeded numpy.random gives the reference's values."""
        rng = np.random if rng is None else rng
        n = self.ns * self.nu * self.nv
        f = np.maximum(0.0, rng.normal(1.0, std, n)).reshape(self.ns, self.nu, self.nv)
        self._set_mat(self._host_mat() * f)

    def add_abs_noise(self, std=0.1, rng=None):
        """Additive N(0, std) noise on the used rays (tomo3D.py:344-352), drawn
        in (i, j, k) order over the used rays only."""
        rng = np.random if rng is None else rng
        used = self._used()
        # float32 value + float64 draw, rounded to float32, as the reference's
        # scalar update does
        m = np.array(self._host_mat(), np.float64)
        m[used] += rng.normal(0.0, std, int(used.sum()))
        self._set_mat(m)

    def add_rel_noise(self, delta=0.002, rng=None):
        """Additive N(0, delta(1 + value/2)) noise on the used rays
        (tomo3D.py:356-366), drawn in (i, j, k) order over the used rays."""
        rng = np.random if rng is None else rng
        used = self._used()
        m = np.array(self._host_mat(), np.float64)
        m[used] += rng.normal(0.0, delta * (1.0 + m[used] / 2.0))
        self._set_mat(m)

    def add_poisson_noise(self, photon_num_per_bin=1000, seed=None):
        """Poisson counts I ~ P(N e^-p), p_hat = -ln(max(I,1)/N).  The reference
        uses an unseeded Gaussian approximation (emil_code/tomo3D.py:368-382);
        this draws true Poisson counts from a seeded generator."""
        rng = np.random.default_rng(seed)
        m = self.mat.cpu().numpy() if _is_tensor(self.mat) else self.mat
        I = rng.poisson(photon_num_per_bin * np.exp(-np.asarray(m, np.float64)))
        p = -np.log(np.maximum(I, 1) / photon_num_per_bin).astype(default_type)
        if _is_tensor(self.mat):
            self.mat = _torch().from_numpy(p).to(self.mat.device)
        else:
            self.mat = p

    def homoscale(self):
        """Divides each ray by its length (tomo3D.py:395-399)."""
        from . import tomo3Dfort
        self.mat = tomo3Dfort.sinoscale(self.mat, self.frame_vectors, self.ns, self.nu, self.nv,
                                        self.du, self.dv, self.u0, self.v0)

    def backproject_to(self, image, alg='adj_linear'):
        """emil_code/tomo3D.py:432-472."""
        bp = _pick(bp_algs_dict, alg, "backprojection")
        image.mat = bp(self.mat, self.indicator, self.frame_vectors, self.ns, self.nu, self.nv,
                       self.du, self.dv, self.u0, self.v0, image.dx, image.dy, image.dz,
                       image.x0, image.y0, image.z0, image.nx, image.ny, image.nz)


class raybunch(sinogram3D):
    """A coarse sinogram of averaged ray windows (tomo3D.py:484-534): coarse bin
    (i, j, k) is the mean of the fine window [i*nds, i*nds+nws) x [j*ndu,
    j*ndu+nwu) x [k*ndv, k*ndv+nwv)."""

    def __init__(self, sino, nds, ndu, ndv, nws, nwu, nwv):
        self.nviews = sino.ns
        self.nds, self.ndu, self.ndv = nds, ndu, ndv
        self.nws, self.nwu, self.nwv = nws, nwu, nwv
        sinogram3D.__init__(self, config_name=sino.config_name, parms=sino.parms,
                            shape=(sino.ns // nds, sino.nu // ndu, sino.nv // ndv),
                            slen=sino.slen, ulen=sino.ulen, vlen=sino.vlen, s0=sino.s0,
                            u0=sino.u0, v0=sino.v0, s_include_endpoints=sino.s_include_endpoints)

    def _n9(self):
        return (self.ns, self.nu, self.nv, self.nds, self.ndu, self.ndv,
                self.nws, self.nwu, self.nwv)

    def get_data(self, sino):
        from . import tomo3Dfort
        self.mat = tomo3Dfort.sinoave(sino.mat, sino.ns, sino.nu, sino.nv, *self._n9())

    def put_data(self, sino):
        from . import tomo3Dfort
        sino.mat = tomo3Dfort.sinodist(self.mat, sino.ns, sino.nu, sino.nv, *self._n9())


class linogram3D:
    """Ray-list data (emil_code/tomo3D.py:541-645)."""

    def __init__(self, config_name='ball', parms={"source_radius": 1., "det_radius": 1.},
                 nrays=default_num_rays, is_duplicate=0):
        self.mat = np.zeros([nrays], default_type)
        self.config_name = config_name
        self.parms = parms
        self.nr = nrays
        self.frame = configs3D.lino_configs.get(config_name)
        if not is_duplicate:
            self.ray_vectors = np.zeros((self.nr, 6), default_type)
            self.indicator = np.ones([nrays], indicator_type)

    def duplicate(self):
        new = linogram3D(self.config_name, dict(self.parms), self.nr, is_duplicate=1)
        new.mat = self.mat.clone() if _is_tensor(self.mat) else self.mat.copy()
        new.indicator = self.indicator
        new.ray_vectors = self.ray_vectors
        return new

    def frame_vecs(self, s):
        c = self.frame(s, self.parms)
        return np.array(c[0:3]), np.array(c[3:6])

    def dist_to(self, lino):
        return _xp_sum_sq(self.mat, lino.mat)

    def copy_to(self, dest):
        dest.mat = self.mat.clone() if _is_tensor(self.mat) else self.mat.copy()
        dest.indicator = self.indicator.clone() if _is_tensor(self.indicator) else self.indicator.copy()

    def backproject_to(self, image):
        """emil_code/tomo3D.py:620-645: adjoint of nearest ray projection."""
        image.mat = ray_proj3D.sinobackproject_nearest(
            self.mat, self.indicator, self.ray_vectors, self.nr, image.dx, image.dy, image.dz,
            image.x0, image.y0, image.z0, image.nx, image.ny, image.nz)


class image3D:
    """3D image array (emil_code/tomo3D.py:784-1500)."""

    def __init__(self, shape=default_image_shape, xlen=2.0, ylen=2.0, zlen=2.0,
                 x0=-1.0, y0=-1.0, z0=-1.0, data_type='real'):
        self.mat = np.zeros(shape, default_type)
        self.indicator = np.ones(shape, indicator_type)
        self.nx, self.ny, self.nz = shape
        self.x0, self.y0, self.z0 = x0, y0, z0
        self.xlen, self.ylen, self.zlen = xlen, ylen, zlen
        self.dx = self.get_dx()
        self.dy = self.get_dy()
        self.dz = self.get_dz()

    def get_dx(self):
        return self.xlen / self.nx

    def get_dy(self):
        return self.ylen / self.ny

    def get_dz(self):
        return self.zlen / self.nz

    def add_shape(self, shape):
        """tomo3D.py:1073-1085: adds shape.voxvalf at every voxel centre to
        mat; one launch over the volume instead of a Python loop over voxels."""
        from . import quadric
        quadric._embed_image(shape._voxval_tables(), self)

    def _grid(self):
        return (self.dx, self.dy, self.dz, self.x0, self.y0, self.z0, self.nx, self.ny, self.nz)

    def read(self, file_name, data_type=default_type):
        """tomo3D.py:202-208 / :876-880: raw array of the current shape; a
        device array stays on the device."""
        shape = tuple(self.mat.shape)
        a = np.fromfile(file_name, dtype=data_type).reshape(shape)
        if _is_tensor(self.mat):
            self.mat = _torch().from_numpy(a).to(self.mat.device)
        else:
            self.mat = a

    def write_file(self, file_name):
        """tomo3D.py:254-257 / :883-885: raw native bytes in C order."""
        from . import Utils
        Utils.write_dat(file_name, self.mat)

    def duplicate(self):
        new = image3D((self.nx, self.ny, self.nz), self.xlen, self.ylen, self.zlen,
                      self.x0, self.y0, self.z0)
        new.mat = self.mat.clone() if _is_tensor(self.mat) else self.mat.copy()
        return new

    def to_device(self):
        torch = _torch()
        from . import _dev
        self.mat = _dev.to_dev(self.mat, torch.float32, (self.nx, self.ny, self.nz))
        return self

    def copy_to(self, dest):
        dest.mat = self.mat.clone() if _is_tensor(self.mat) else self.mat.copy()

    def clear(self):
        self.mat *= 0.

    def fill(self, v):
        self.mat[...] = v

    def make_positive(self):
        """emil_code/tomo3D.py:902-903."""
        if _is_tensor(self.mat):
            self.mat.clamp_(min=0.0)
        else:
            self.mat[self.mat < 0.] = 0.

    def make_positive_softly(self, beta=1.0):
        self.mat[self.mat < 0.] *= (1.0 - beta)

    def zero_edges(self):
        """Two-voxel border set to zero (emil_code/tomo3D.py:928-943)."""
        m = self.mat
        for sl in (np.s_[0:2, :, :], np.s_[-2:, :, :], np.s_[:, 0:2, :], np.s_[:, -2:, :],
                   np.s_[:, :, 0:2], np.s_[:, :, -2:]):
            m[sl] = 0.

    def clamp_edges(self):
        """emil_code/tomo3D.py:907-925."""
        m = self.mat
        m[0], m[1], m[-1], m[-2] = m[2] * 1., m[2] * 1., m[-3] * 1., m[-3] * 1.
        m[:, 0], m[:, 1], m[:, -1], m[:, -2] = m[:, 2] * 1., m[:, 2] * 1., m[:, -3] * 1., m[:, -3] * 1.
        m[:, :, 0], m[:, :, 1] = m[:, :, 2] * 1., m[:, :, 2] * 1.
        m[:, :, -1], m[:, :, -2] = m[:, :, -3] * 1., m[:, :, -3] * 1.

    def dist_to(self, image):
        """emil_code/tomo3D.py:946-951 (float64)."""
        return _xp_sum_sq(self.mat, image.mat)

    def calc_total_var(self, lam1=0.):
        """emil_code/tomo3D.py:953-958."""
        return float(totalvar.total_var3d(lam1, self.mat, self.nx, self.ny, self.nz))

    def calc_total_var_p(self, p=1.):
        """emil_code/tomo3D.py:960-965."""
        return float(totalvar.total_var3d_p(p, self.mat, self.nx, self.ny, self.nz))

    def calc_lp(self, p=1.):
        """emil_code/tomo3D.py:967-972."""
        return float(totalvar.calc_lp(p, self.mat, self.nx, self.ny, self.nz))

    def calc_total_dvar_p(self, p=1.):
        """emil_code/tomo3D.py:975-980."""
        return float(totalvar.total_dvar3d_p(p, self.mat, self.nx, self.ny, self.nz,
                                             self.dx, self.dy, self.dz))

    def lp_grad(self, p=1., normalizeq=1):
        """emil_code/tomo3D.py:982-987: mat <- gradient, returns its norm."""
        self.mat, bobo = totalvar.lp_grad3d(p, self.mat, self.nx, self.ny, self.nz, normalizeq)
        return bobo

    def tpv_grad(self, p=1., normalizeq=1):
        """emil_code/tomo3D.py:990-995."""
        self.mat, bobo = totalvar.total_p_var_grad3d(p, self.mat, self.nx, self.ny, self.nz,
                                                     normalizeq)
        return bobo

    def tpdv_grad(self, p=1., normalizeq=1):
        """emil_code/tomo3D.py:999-1004."""
        self.mat, bobo = totalvar.total_p_dvar_grad3d(p, self.mat, self.nx, self.ny, self.nz,
                                                      self.dx, self.dy, self.dz, normalizeq)
        return bobo

    def gmi(self):
        """emil_code/tomo3D.py:1025-1028: mat <- gradient magnitude image."""
        self.mat = totalvar.gmi(self.mat, self.nx, self.ny, self.nz)

    def tv_grad(self, lam1=0., normalizeq=1):
        """emil_code/tomo3D.py:1031-1036: mat <- TV gradient, returns its norm."""
        self.mat, bobo = totalvar.total_var_grad3d(lam1, self.mat, self.nx, self.ny, self.nz,
                                                   normalizeq)
        return bobo

    def scale(self, factor):
        self.mat = factor * self.mat

    # weighted ray-bunch and derivative ART (tests/cweighted_art_test.py,
    # tests/cderiv_art_test.py); launch-shape arguments are ignored
    def cweighted_project_to(self, data, rayweights, nss, nsu, nsv, *launch):
        from . import weighted
        weighted.cweighted_project_to(self, data, rayweights, nss, nsu, nsv)

    def put_weights(self, data, rayweights, nss, nsu, nsv, *launch):
        from . import weighted
        weighted.put_weights(self, data, rayweights, nss, nsu, nsv)

    def c_ss_weighted_art_step(self, data, rayweights, nss, nsu, nsv, *launch, reg_fac=1.0):
        from . import weighted
        weighted.c_ss_weighted_art_step(self, data, rayweights, nss, nsu, nsv, reg_fac=reg_fac)

    def get_deriv_weights(self, data, *launch):
        from . import weighted
        weighted.get_deriv_weights(self, data)

    def cderiv_art_step(self, data, *launch, reg_fac=1.0):
        from . import weighted
        weighted.cderiv_art_step(self, data, reg_fac=reg_fac)

    def project_to(self, sino, alg='proj_linear'):
        """emil_code/tomo3D.py:1127-1168."""
        proj = _pick(proj_algs_dict, alg, "projection")
        sino.mat = proj(sino.indicator, sino.frame_vectors, sino.ns, sino.nu, sino.nv, sino.du,
                        sino.dv, sino.u0, sino.v0, self.mat, *self._grid())

    def lino_project_to(self, lino, alg='proj_linear'):
        """emil_code/tomo3D.py:1172-1195."""
        proj = _pick(ray_proj_algs_dict, alg, "projection")
        lino.mat = proj(lino.indicator, lino.ray_vectors, lino.nr, self.mat, *self._grid())

    def art_step(self, sino, reg_fac=1.0, alg='art_linear'):
        """emil_code/tomo3D.py:1235-1277."""
        art = _pick(art_algs_dict, alg, "ART")
        self.mat = art(reg_fac, sino.mat, sino.indicator, sino.frame_vectors, sino.ns, sino.nu,
                       sino.nv, sino.du, sino.dv, sino.u0, sino.v0, self.mat, *self._grid())

    def tomo_art_step_2vol(self, iaux, sino, reg_fac=1.0):
        """emil_code/tomo3D.py:1282-1325."""
        self.mat, iaux.mat = vpt.artstepz2vol(
            reg_fac, sino.mat, sino.indicator, sino.frame_vectors, sino.ns, sino.nu, sino.nv,
            sino.du, sino.dv, sino.u0, sino.v0, self.mat, *self._grid(), iaux.mat, *iaux._grid())

    def lino_art_step(self, lino, reg_fac=1.0, alg='art_linear'):
        """emil_code/tomo3D.py:1329-1353."""
        art = _pick(ray_art_algs_dict, alg, "ART")
        self.mat = art(reg_fac, lino.mat, lino.indicator, lino.ray_vectors, lino.nr, self.mat,
                       *self._grid())

    def em_step(self, sino, alg='em_nearest'):
        """emil_code/tomo3D.py:1394-1435."""
        em = _pick(em_algs_dict, alg, "EM")
        self.mat = em(sino.mat, sino.indicator, sino.frame_vectors, sino.ns, sino.nu, sino.nv,
                      sino.du, sino.dv, sino.u0, sino.v0, self.mat, *self._grid())

    def lino_em_step(self, lino):
        """emil_code/tomo3D.py:1476-1499."""
        self.mat = ray_proj3D.emstep_nearest(lino.mat, lino.indicator, lino.ray_vectors, lino.nr,
                                             self.mat, *self._grid())
