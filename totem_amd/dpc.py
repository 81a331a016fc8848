"""Differential phase contrast from segmented-detector STEM images (``ToTEM_submit.m:1009-1128``).

The scan images of the four quadrants of a ring (masks from :func:`optics.dpc_masks`, detector
sums from the engine) give the DPC vector image px = B - D, py = C - A, and integrated DPC
p = Re IFFT2(ifftshift((F(px) kx + F(py) ky) / (2 pi i |k|^2))), F = fftshift(FFT2), with the
zero frequency set to mean(px) + mean(py), on the scan grid kx = (-W/2 .. W/2-1) / (W sampling step).
This is the reference's post-processing of the detector images on the host.
"""
import numpy as np


def scan_image(myresul_column, width_red, hight_red, norm=1.0):
    """One detector column of myresul [D] as the scan image [hight_red, width_red] (row = scan row),
    divided by ``norm`` (the reference's probesx^2 * allvib)."""
    return np.asarray(myresul_column, dtype=np.float64).reshape(hight_red, width_red) / norm


def ddpc(myresul, width_red, hight_red, ring=0, norm=1.0):
    """(px, py, sum) of ring ``ring``: quadrant images A, B, C, D are detector columns 4*ring .. 4*ring+3."""
    a, b, c, d = (scan_image(myresul[:, 4 * ring + q], width_red, hight_red, norm) for q in range(4))
    return b - d, c - a, a + b + c + d


def idpc(px, py, sampling_nm, step):
    """Integrated DPC image from the vector image (``ToTEM_submit.m:1012,1081-1083``)."""
    h, w = px.shape
    if h % 2 or w % 2:
        raise ValueError("IDPC needs even scan sizes (the k grid (-W/2 .. W/2-1) centres zero at W/2)")
    kx, ky = np.meshgrid(np.arange(-w // 2, w // 2) / (w * sampling_nm * step),
                         np.arange(-h // 2, h // 2) / (h * sampling_nm * step))
    k2 = kx * kx + ky * ky
    num = np.fft.fftshift(np.fft.fft2(px)) * kx + np.fft.fftshift(np.fft.fft2(py)) * ky
    with np.errstate(divide="ignore", invalid="ignore"):
        p = num / (2 * np.pi * 1j * k2)
    p[h // 2, w // 2] = px.mean() + py.mean()
    return np.real(np.fft.ifft2(np.fft.ifftshift(p)))
