"""Plane-wave HRTEM and CBED multislice on the engine (``ToTEM_submit.m:1185-1473``).

Both paths take the transmission stack of the potential-only entry point (built on the GPU, as the
reference's GPU branch does through ``cuda_stem_potential``) and run the multislice loop in the
engine's STEM kernels with one work item per wave:

* HRTEM: a (tilted) plane wave over the whole potential, propagator cut at 1/2 of the Nyquist
  radius; the exit wave Psi = FFT2(psi) comes out of the last columns pass; the image
  sum_m gfsf_m |IFFT2(ifftshift(Psi tcc_m ap))|^2 is formed by ``totem_tcc_image``.
* CBED: the probe stack at one window of the potential, propagator cut at 2/3; the pattern
  sum_m gfsf_m |fftshift(Psi_m)|^2 by ``totem_weighted_pattern``.
"""
import numpy as np

from . import engine, optics


def plane_wave(n, tilt_pixels=(0, 0)):
    """The reference's incident wave: ifft2(ifftshift(delta)) * n^2, delta at the centre minus the
    tilt in pixels (x, y) (``ToTEM_submit.m:1305-1311``)."""
    f = np.zeros((n, n), dtype=np.complex128)
    c = n // 2 if n % 2 == 0 else (n - 1) // 2
    f[c - int(round(tilt_pixels[1])), c - int(round(tilt_pixels[0]))] = 1.0
    return np.fft.ifft2(np.fft.ifftshift(f)) * n * n


def _exit_waves(t_dev, np_side, slices, waves, prop, r0, c0, wrap):
    """Exit waves Psi_m = FFT2(psi_m) (FFT order) of the given incident waves [M, n, n] through the
    first `slices` slices of the device transmission stack, window at (r0, c0); a device tensor."""
    import torch
    M, n = waves.shape[0], waves.shape[-1]
    probes = torch.tensor(np.asarray(waves).astype(np.complex64), device="cuda")
    h = torch.tensor(optics.unit_complex64(prop), device="cuda")
    ones = torch.ones((1, n, n), dtype=torch.float64, device="cuda")      # one (unused) detector
    det = torch.zeros((1, 1, M), dtype=torch.float64, device="cuda")
    out = torch.zeros((M, n, n), dtype=torch.complex64, device="cuda")
    job = engine.stem_job(n, np_side, slices, t_dev.data_ptr(), probes.data_ptr(), M, h.data_ptr(), ones.data_ptr(),
                          1, 1, 1, 0, r0, c0, det.data_ptr(), wrap=wrap, exit_ptr=out.data_ptr())
    engine.stem_run(job, M, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return out


def hrtem(t_dev, slices, dx, kv, slice_thickness, tcc, gfsf, ap, tilt_mrad=0.0, tilt_phi_deg=0.0):
    """(image [n, n], Psi centred [n, n] complex128) of a plane wave through the n x n transmission
    stack t_dev (device complex64 [S, n, n])."""
    n = int(t_dev.shape[-1])
    lam = optics.wavelength_nm(kv)
    tilt = tilt_mrad * 1e-3
    phi = -tilt_phi_deg
    sampling_nm = dx * 0.1
    tiltxnum = tilt / lam * np.cos(phi / 180 * np.pi) * sampling_nm * n
    tiltynum = tilt / lam * np.sin(phi / 180 * np.pi) * sampling_nm * n
    wave0 = plane_wave(n, (tiltxnum, tiltynum) if tilt > 0 else (0, 0))
    prop = optics.propagator(n, dx, kv, slice_thickness, tilt_mrad, tilt_phi_deg, cut=16.0)
    psi_k = _exit_waves(t_dev, n, slices, wave0[None], prop, 0, 0, True)
    Psi = np.fft.fftshift(psi_k[0].cpu().numpy().astype(np.complex128))
    return engine.tcc_image(Psi, tcc, ap, gfsf), Psi


def cbed(t_dev, np_side, slices, dx, kv, slice_thickness, probes, gfsf, r0, c0):
    """CBED pattern [n, n] of the probe stack [M, n, n] at window origin (r0, c0) (0-based)."""
    n = probes.shape[-1]
    prop = optics.propagator(n, dx, kv, slice_thickness)
    psi_k = _exit_waves(t_dev, np_side, slices, probes, prop, r0, c0, False)
    return engine.weighted_pattern(psi_k.data_ptr(), n, gfsf, shift=True)
