"""Python-3 restatement of the reference's FDK (emil_code/FDK.py) -- TEST
INFRASTRUCTURE ONLY (the checker of ct2.FDK; never imported by the product).

Float64 numpy, the reference's own route: the pre-weight, the zero-padded
FFT ramp filter with numpy.fft (pocketfft) and the voxel-driven bilinear
backprojection.  Python-2 idioms of the original (integer `/`, `print`
statements, n.integer) are the only changes.
"""
from __future__ import annotations

import numpy as np


def fdk_sino_weight3D(radius, center2detector, ns, vu, vv, sinomat):
    """FDK.py:10-23."""
    F = radius
    D = center2detector
    weight = F / np.sqrt((F + D) ** 2 + vv ** 2 + vu ** 2)
    return np.tile(weight, [ns, 1, 1]) * sinomat


def kappa1D(N, delta):
    """setFreqrange.py:3-10."""
    if N % 2 == 0:
        return np.arange(-N // 2, N // 2, dtype=float) / (delta * N)
    return np.arange(-(N - 1) // 2, (N - 1) // 2 + 1, dtype=float) / (delta * N)


def ramp_filter(dp, n, opt_window='', L=1):
    """Filt of FDK.py:33-73 for lines of n samples; returns (npad, Filt)."""
    npad = int(2 ** (np.round(np.log2(2 * n - 1))))
    N = npad
    if N % 2 == 0:
        nn = np.arange(-N / 2, N / 2, dtype=float)
    else:
        nn = np.arange(-(N - 1) / 2, (N - 1) / 2 + 1, dtype=float)
    hn = np.zeros(nn.size)
    hn[npad // 2] = 1. / (4 * dp ** 2)
    for ii in range(hn.size):
        if abs(nn[ii]) % 2 == 1:
            hn[ii] = -1. / (np.pi * nn[ii] * dp) ** 2
    hn = np.fft.ifftshift(hn)
    Hk = np.real(np.fft.fft(hn))
    ramp = dp * Hk
    if opt_window == 'Hann':
        kappa = kappa1D(npad, dp)
        cutoff_freq = 1 / (2 * dp) * L
        if L < 1.0:
            diff = np.abs(np.abs(kappa) - cutoff_freq)
            mindiff = diff.min()
            kappaindex = np.nonzero(diff == mindiff)[0]
            startindex, endindex = kappaindex[0], kappaindex[-1]
            HannFilt = np.zeros(kappa.shape)
            HannFilt[startindex:endindex + 1] = 0.5 + 0.5 * np.cos(
                np.pi * kappa[startindex:endindex + 1] / cutoff_freq)
            HannFilt = np.fft.ifftshift(HannFilt)
        else:
            HannFilt = np.fft.ifftshift(0.5 + 0.5 * np.cos(np.pi * kappa / cutoff_freq))
        return npad, ramp * HannFilt
    return npad, ramp


def fdk_sino_filter3D(dp, dim, sinomat, opt_window='', L=1):
    """FDK.py:25-103 (the per-line Python loop vectorised over lines with
    numpy.fft along the filtered axis -- the same transforms)."""
    origsize = sinomat.shape[dim]
    npad, Filt = ramp_filter(dp, origsize, opt_window, L)
    pad = [(0, 0)] * 3
    pad[dim] = (0, npad - origsize)
    x = np.pad(np.asarray(sinomat, np.float64), pad)
    shape = [1, 1, 1]
    shape[dim] = npad
    y = np.real(np.fft.ifft(np.fft.fft(x, axis=dim) * Filt.reshape(shape), axis=dim))
    sl = [slice(None)] * 3
    sl[dim] = slice(0, origsize)
    return y[tuple(sl)]


def fdk_backproj_linear3D(ns, nu, nv, du, dv, u0, v0, radius, source2detector, uoffset, voffset,
                          vs, vx, vy, vz, sinomat):
    """FDK.py:130-208; returns recon[vx.size, vz.size] scaled by pi/ns."""
    F = radius
    D = source2detector - radius
    recon = np.zeros((vx.size, vz.size))
    for jj in range(vz.size):
        for ii in range(ns):
            zoomfac = (F + D) / (F - vx * np.cos(vs[ii]) - vy * np.sin(vs[ii]))
            uinlen = (-vx * np.sin(vs[ii]) + vy * np.cos(vs[ii])) * zoomfac
            vinlen = vz[jj] * zoomfac
            lu = np.floor((uinlen - u0) / du - uoffset).astype(np.int64)
            lv = np.floor((vinlen - v0) / dv - voffset).astype(np.int64)
            tmpsino = sinomat[ii, :, :]
            mapping = (lu + 1 < nu) * (lu >= 0) * (lv + 1 < nv) * (lv >= 0)
            m = np.nonzero(mapping)
            a11 = tmpsino[lu[m], lv[m]]
            a21 = tmpsino[lu[m] + 1, lv[m]]
            a12 = tmpsino[lu[m], lv[m] + 1]
            a22 = tmpsino[lu[m] + 1, lv[m] + 1]
            wru = (uinlen[m] - u0) / du - (lu[m] + uoffset)
            wlu = 1 - wru
            wrv = (vinlen[m] - v0) / dv - (lv[m] + voffset)
            wlv = 1 - wrv
            v1 = wlu * a11 + wru * a21
            v2 = wlu * a12 + wru * a22
            val = wlv * v1 + wrv * v2
            recon[m, jj] = recon[m, jj] + val * zoomfac[m] * zoomfac[m]
    return np.pi / ns * recon


def recon_cube(recon, nx, ny, nz):
    """recon3D.py:134-139: recon[ny*nx, nz] -> cube[nx, ny, nz]."""
    cube = np.zeros((nx, ny, nz))
    for ii in range(nz):
        cube[:, :, ii] = np.reshape(recon[:, ii], (ny, nx)).T
    return cube
