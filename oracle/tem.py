"""FP64 restatement of the HRTEM and CBED multislice paths of ``ToTEM_submit.m`` (oracle; test
infrastructure only).

HRTEM (``HRTEMsimulation``, ``ToTEM_submit.m:1185-1341``): a plane wave (``:1305-1311``) through every
slice -- psi = ifft2(fft2(psi .* t_k) .* H) with H cut at |k|^2 > 1/(16 dx^2) (``:1298-1302``,
``:1314-1320``) -- then Psi = fftshift(fft2(psi)) (``:1333``) and
image = sum_m gfsf_m |ifft2(ifftshift(Psi .* tcc_m .* ap))|^2 (``:1336-1338``).
CBED (``CBEDsimulation``, ``:1344-1473``): for every focal member the probe through the slices cropped
at the CBED window with H cut at 2/3 (``:1374-1378``), pattern = sum_m gfsf_m |fftshift(fft2(psi_m))|^2
(``:1457-1469``).

Pins (``tests/test_tem.py``): a vacuum specimen gives a uniform HRTEM image of value sum gfsf and a
CBED pattern equal to the probes' band-limited diffraction intensity.
"""
import numpy as np


def plane_wave(n, tilt_pixels=(0, 0)):
    """ifft2(ifftshift(delta)) * n^2 with the delta at the centre minus the tilt (``:1305-1311``)."""
    f = np.zeros((n, n), dtype=np.complex128)
    c = n // 2 if n % 2 == 0 else (n - 1) // 2
    f[c - int(round(tilt_pixels[1])), c - int(round(tilt_pixels[0]))] = 1.0
    return np.fft.ifft2(np.fft.ifftshift(f)) * n * n


def hrtem(t, psf_fft, wave0, tcc, gfsf, ap):
    """(image [n, n], exit wave psi [n, n], Psi centred [n, n])."""
    psi = np.asarray(wave0, dtype=np.complex128)
    for k in range(t.shape[0]):
        psi = np.fft.ifft2(np.fft.fft2(psi * t[k]) * psf_fft)
    Psi = np.fft.fftshift(np.fft.fft2(psi))
    img = np.zeros(psi.shape)
    for m in range(len(gfsf)):
        img += gfsf[m] * np.abs(np.fft.ifft2(np.fft.ifftshift(Psi * tcc[m] * ap))) ** 2
    return img, psi, Psi


def cbed(t, probes, gfsf, psf_fft, r0, c0):
    """CBED pattern [n, n] of the probe stack at window origin (r0, c0) (0-based)."""
    n = probes.shape[-1]
    out = np.zeros((n, n))
    for m in range(len(gfsf)):
        psi = np.asarray(probes[m], dtype=np.complex128)
        for k in range(t.shape[0]):
            psi = np.fft.ifft2(np.fft.fft2(psi * t[k, r0:r0 + n, c0:c0 + n]) * psf_fft)
        out += gfsf[m] * np.abs(np.fft.fftshift(np.fft.fft2(psi))) ** 2
    return out
