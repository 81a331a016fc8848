"""Plane-wave HRTEM and CBED multislice (ToTEM_submit.m:1185-1473) on the engine.

CPU: pins of the FP64 oracle (oracle/tem.py) -- a vacuum specimen gives a uniform HRTEM image equal to
sum gfsf and a CBED pattern equal to the probes' band-limited diffraction intensity; the tilted plane
wave is the expected phase ramp.
GPU: the engine paths (potential built by the engine, multislice in the STEM kernels, imaging by
totem_tcc_image / totem_weighted_pattern) against the FP64 oracle on the oracle's potential,
<= 1e-5 relative L2.
"""
import numpy as np
import pytest

from oracle import potential as opot
from oracle import tem as otem
from totem_amd import optics, specimen, stem


def test_hrtem_vacuum_gives_uniform_image():
    n, dx, kv = 64, 0.2, 200.0
    t = np.ones((5, n, n), dtype=np.complex128)
    psf = optics.propagator(n, dx, kv, 1.95, cut=16.0)
    tcc, gfsf, ap = optics.tcc_stack(n, dx, kv, 12.0, optics.Aberrations(C1=-5.0, Cs_um=1.0, A1=(3.0, 20.0)), 3, 3.0,
                                     0.2)
    img, psi, Psi = otem.hrtem(t, psf, otem.plane_wave(n), tcc, gfsf, ap)
    np.testing.assert_allclose(psi, 1.0, atol=1e-12)
    np.testing.assert_allclose(img, gfsf.sum(), rtol=1e-12)


def test_tilted_plane_wave_is_a_phase_ramp():
    n = 32
    w = otem.plane_wave(n, (2, 0))
    col = np.arange(n)
    np.testing.assert_allclose(w, np.exp(-2j * np.pi * 2 * col / n)[None, :].repeat(n, 0), atol=1e-12)


def test_cbed_vacuum_is_band_limited_probe_intensity():
    n, dx, kv = 64, 0.15, 300.0
    probes, gfsf = optics.probe_stack(n, dx, kv, 24.0, optics.Aberrations(C1=-5.0), 3, 3.0)
    t = np.ones((4, 96, 96), dtype=np.complex128)
    psf = optics.propagator(n, dx, kv, 1.95)
    pat = otem.cbed(t, probes, gfsf, psf, 10, 20)
    band = np.fft.fftshift(np.abs(psf) > 0)
    ref = sum(g * np.abs(np.fft.fftshift(np.fft.fft2(p))) ** 2 for g, p in zip(gfsf, probes)) * band
    np.testing.assert_allclose(pat, ref, rtol=1e-10, atol=1e-10 * ref.max())


def _specimen(n, np_side, cells, S, flag="l"):
    a = 3.905
    x = specimen.perovskite(cells, cells, S // 2, a=a)
    rng = np.random.default_rng(11)
    x[:, 1:3] = (x[:, 1:3] + rng.normal(0, 0.05, (len(x), 2))) % (cells * a)
    dz = a / 2
    cfg = stem.StemConfig(atoms=x, np_side=np_side, dx=cells * a / np_side, n=n, kv=200.0, alpha_mrad=20.0,
                          slice_thickness=dz, n_slices=S, slice_origin=-dz / 4, paraflag=flag)
    return cfg, stem.prepare(cfg, optics_too=False)


def _device_t(engine, inp):
    import torch
    c = inp.cfg
    t = torch.empty((c.n_slices, c.np_side, c.np_side), dtype=torch.complex64, device="cuda")
    engine.build_transmission(c.paraflag, c.np_side, c.n_slices, inp.series_n, inp.ele_n, inp.gx, inp.gy, inp.s2,
                              inp.aper_true, inp.sigma, inp.parameter, t.data_ptr(),
                              stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return t


def _oracle_t(inp):
    c = inp.cfg
    _, t = opot.potential_slices(c.np_side, inp.ele_n, None, inp.ele_n_i, None, inp.series_n, inp.series_n_i,
                                 inp.s2, inp.gx, inp.gy, inp.sigma, inp.parameter, inp.aper_true, c.paraflag)
    return t


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.gpu
def test_hrtem_matches_oracle(engine):
    from totem_amd import tem
    print("case hrtem plane wave 128^2 x 8 slices", flush=True)
    cfg, inp = _specimen(128, 128, 2, 8)
    tcc, gfsf, ap = optics.tcc_stack(128, cfg.dx, cfg.kv, 15.0, optics.Aberrations(C1=-8.0, Cs_um=1.2), 3, 4.0, 0.3)
    img, Psi = tem.hrtem(_device_t(engine, inp), cfg.n_slices, cfg.dx, cfg.kv, cfg.slice_thickness, tcc, gfsf, ap)
    psf = optics.propagator(128, cfg.dx, cfg.kv, cfg.slice_thickness, cut=16.0)
    img_o, _, Psi_o = otem.hrtem(_oracle_t(inp), psf, otem.plane_wave(128), tcc, gfsf, ap)
    e_img, e_psi = _rel(img, img_o), _rel(np.abs(Psi) ** 2, np.abs(Psi_o) ** 2)
    print("  image rel L2 %.3e, exit-wave intensity rel L2 %.3e (FP64-oracle error)" % (e_img, e_psi))
    assert e_img <= 1e-5 and e_psi <= 1e-5


@pytest.mark.gpu
def test_cbed_matches_oracle(engine):
    from totem_amd import tem
    print("case cbed 128^2 window of 256^2 x 8 slices", flush=True)
    cfg, inp = _specimen(128, 256, 4, 8)
    probes, gfsf = optics.probe_stack(128, cfg.dx, cfg.kv, 20.0, optics.Aberrations(C1=-4.0), 2, 3.0)
    pat = tem.cbed(_device_t(engine, inp), 256, cfg.n_slices, cfg.dx, cfg.kv, cfg.slice_thickness, probes, gfsf,
                   37, 61)
    psf = optics.propagator(128, cfg.dx, cfg.kv, cfg.slice_thickness)
    ref = otem.cbed(_oracle_t(inp), probes, gfsf, psf, 37, 61)
    e = _rel(pat, ref)
    print("  pattern rel L2 %.3e (FP64-oracle error)" % e)
    assert e <= 1e-5
