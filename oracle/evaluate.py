"""FP64 CPU restatements of the evaluation against the forward model (test
infrastructure only): the band-limited ground-truth potential of an atom set
and the multislice reprojection of a volume with the reference's error
measure.  The GPU forms are utr_ground_truth_device and
utr_reproject_device (include/utr_hip.h).

ground_truth: the potential of vatom's parameterisation
(slicelib.cpp:2699-2737), the function autosliccmd's noutput 3 samples
(autosliccmd.cpp:1146-1151), restricted to |q| < qmax on the periodic cube of
side L = n xst and evaluated exactly by Fourier synthesis with
Vt(q) = (al/pi) sum a_i/(q^2 + b_i) + ag pi^-1.5 sum c_i exp(-d_i q^2).

reproject: DHT.cpp:1509-1731 -- the volume times pi/(lambda E) zst
(DHT.cpp:1514) is turned to each orientation (about Z, then Y', as
autosliccmd.cpp:709-748 turns atoms), sampled trilinearly, and propagated
plane by plane with the multislice of oracle/msct.py (one slice per plane,
bandwidth limit, propagator of thickness zst); the exit wave goes to the
defocus distance (R = defocus - L/2, as autosliccmd.cpp:1065) and the error of
DHT.cpp:1686 is || |psi| - sqrt(I0) ||^2 / || sqrt(I0) ||^2 with I0
renormalised and clamped (DHT.cpp:1641-1647).
"""
import math

import numpy as np
from scipy import ndimage

from . import msct as M

AL, AG = 150.4121417, 266.5985798


def vtilde(Z, q2):
    """3D Fourier transform (V A^3) of vatom for element Z at q^2 (A^-2)."""
    fp = M.fe_params()[Z]
    s = sum(AL / math.pi * fp[i] / (q2 + fp[i + 1]) for i in (0, 2, 4))
    return s + sum(AG * math.pi ** -1.5 * fp[i] * np.exp(-fp[i + 1] * q2) for i in (6, 8, 10))


def ground_truth(n, xst, Z, x, y, z, qmax, scale=1.0):
    """[n][n][n] potential (V times scale) at voxel (i, j, k) = cube position
    (i, j, k) xst; atoms in cube coordinates (A).  qmax <= 0: no band limit."""
    L = n * xst
    k = np.fft.fftfreq(n, 1.0 / n)
    KZ, KY, KX = np.meshgrid(k, k, k, indexing="ij")
    q2 = (KX ** 2 + KY ** 2 + KZ ** 2) / L ** 2
    S = np.zeros((n, n, n), np.complex128)
    Z = np.asarray(Z)
    for zv in np.unique(Z):
        sel = Z == zv
        F = np.zeros_like(S)
        for xa, ya, za in zip(np.asarray(x, np.float64)[sel], np.asarray(y, np.float64)[sel],
                              np.asarray(z, np.float64)[sel]):
            F += np.exp(-2j * np.pi * (KX * xa + KY * ya + KZ * za) / L)
        S += vtilde(int(zv), q2) * F
    if qmax > 0:
        S[q2 >= qmax * qmax] = 0
    return np.fft.ifftn(S).real * (n ** 3 / L ** 3) * scale


def electron_ev(wl):
    """Electron energy (eV) from the wavelength (A), UTR.cpp:353-363."""
    hp, cc, ee, m0 = 6.62607004e-34, 299792458.0, 1.602176634e-19, 9.1093837015e-31
    mc2 = m0 * cc * cc
    chl2 = (cc * cc * hp * hp) / (wl * wl * 1e-20)
    return (-mc2 + math.sqrt(mc2 * mc2 + chl2)) / ee


def turn_matrix(angZ_deg, angY_deg):
    """R = Ry(angY) Rz(angZ) of autosliccmd.cpp:709-748 (atoms p -> c + R (p - c))."""
    a, b = math.radians(angZ_deg), math.radians(angY_deg)
    cZ, sZ, cY, sY = math.cos(a), math.sin(a), math.cos(b), math.sin(b)
    return np.array([[cY * cZ, -cY * sZ, sY], [sZ, cZ, 0.0], [-sY * cZ, sY * sZ, cY]])


def turned(vol, angZ_deg, angY_deg, centre):
    """vol sampled on the turned frame: V'(p') = V(c + R^T (p' - c)),
    trilinear with the volume extended by zeros (a sample within one voxel
    outside interpolates toward zero); vol [z][y][x], centre (cx, cy, cz) in
    voxels."""
    nz, ny, nx = vol.shape
    R = turn_matrix(angZ_deg, angY_deg)
    Kz, Ky, Kx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    d = np.stack([Kx - centre[0], Ky - centre[1], Kz - centre[2]])
    src = np.tensordot(R.T, d, axes=1)
    sx, sy, sz = src[0] + centre[0], src[1] + centre[1], src[2] + centre[2]
    return ndimage.map_coordinates(np.asarray(vol, np.float64), [sz, sy, sx], order=1, mode="grid-constant", cval=0.0)


def reproject(p, vol, angZ, angY, defocus, images=None):
    """(intensities [rows][ny][nx], errors [rows] or None) of a UTR volume
    (UTR units) through the multislice; p: UTR parameter dict (nx = ny = nz)."""
    n, xst, wl = p["nx"], p["xst"], p["wl"]
    E = electron_ev(wl)
    kev = E / 1000.0
    L = n * xst
    m = M.Model(n, n, L, xst, kev)
    phase3 = math.pi / (wl * E) * xst
    centre = (n / 2.0 + p.get("dxc", 0.0) / xst, n / 2.0 + p.get("dyc", 0.0) / xst, n / 2.0 + p.get("dzc", 0.0) / xst)
    aobj = float(np.float32(abs(np.float32(p.get("aobj_mrad", 0.0)) * np.float32(0.001))))
    out = np.zeros((len(angZ), n, n))
    errs = np.zeros(len(angZ)) if images is not None else None
    for r in range(len(angZ)):
        tv = turned(vol, angZ[r], angY[r], centre)
        wave = np.ones((n, n), complex)
        for s in range(n):
            t = np.exp(1j * phase3 * tv[s])
            T = M.cf_fft(t)
            T[~m.keep] = 0
            wave = wave * M.cf_ifft(T)
            W = M.cf_fft(wave) * m.propy[:, None] * m.propx[None, :]
            W[~m.keep] = 0
            wave = M.cf_ifft(W)
        c = m.fresnel(wave, defocus[r] - 0.5 * L, aobj, p.get("cs3_mm", 0.0), p.get("cs5_mm", 0.0))
        out[r] = np.abs(c) ** 2
        if images is not None:
            i0 = np.maximum(np.asarray(images[r], np.float64) / p.get("norm_f1", 1.0) + p.get("norm_f2", 0.0), 0.0)
            errs[r] = np.sum((np.sqrt(np.maximum(out[r], 0)) - np.sqrt(i0)) ** 2) / np.sum(i0)
    return out, errs


def utr_unit_factor(wl):
    """UTR's output is the potential in V times sigma_rel lambda E / pi (its
    scale EE lambda / (2 pi zst) of UTR.cpp:1275 with the non-relativistic
    delta = V / 2E); ground truth compared with UTR volumes carries it."""
    E = electron_ev(wl)
    return (M.sigma(E / 1000.0) / 1000.0) * wl * E / math.pi
