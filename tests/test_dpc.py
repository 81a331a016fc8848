"""Segmented detectors and differential phase contrast (ToTEM_submit.m:1009-1183).

CPU: the four quadrant masks of a ring partition it and are the reference's angular sectors; IDPC
integrates exact spectral gradients of a periodic phase back to the phase (minus its mean); the DPC
vector image is B - D, C - A.
GPU: myProjectfun with two 4-quadrant rings against the FP64 oracle (<= 1e-4 relative per pixel),
and the IDPC images of both agree.
"""
import numpy as np
import pytest

from totem_amd import dpc, optics


def test_quadrants_partition_the_ring_and_follow_the_rotation():
    n, dx, kv = 96, 0.08, 200.0
    m = optics.dpc_masks(n, dx, kv, [(10.0, 40.0, 30.0)])
    ring = optics.detector_masks(n, dx, kv, [(10.0, 40.0)])[0]
    np.testing.assert_array_equal(m.sum(axis=0), ring)
    assert np.all(m >= 0)
    f = optics.centred_freq(n, dx)
    px, py = np.meshgrid(f, f)
    theta = np.degrees(np.arctan2(py, px))
    d = 30.0 - 180.0                                  # rotation 30 deg -> sectors from -150 deg
    for q in range(3):
        sector = (theta >= d + 90 * q) & (theta < d + 90 * (q + 1))
        np.testing.assert_array_equal(m[q], ring * sector)


def test_idpc_integrates_gradients_back_to_the_phase():
    h, w, sampling, step = 32, 48, 0.0078, 3
    s = sampling * step
    y, x = np.meshgrid(np.arange(h) * s, np.arange(w) * s, indexing="ij")
    Lx, Ly = w * s, h * s
    phi = (0.3 * np.cos(2 * np.pi * (2 * x / Lx + y / Ly)) + 0.2 * np.sin(2 * np.pi * 3 * y / Ly)
           + 0.1 * np.sin(2 * np.pi * (x / Lx - 4 * y / Ly)) + 1.7)
    kx, ky = np.meshgrid(np.fft.fftfreq(w, s), np.fft.fftfreq(h, s))
    F = np.fft.fft2(phi)
    gx = np.real(np.fft.ifft2(2j * np.pi * kx * F))
    gy = np.real(np.fft.ifft2(2j * np.pi * ky * F))
    np.testing.assert_allclose(dpc.idpc(gx, gy, sampling, step), phi - phi.mean(), atol=1e-12)


def test_idpc_rejects_odd_scan_sizes():
    with pytest.raises(ValueError):
        dpc.idpc(np.zeros((5, 6)), np.zeros((5, 6)), 0.01, 1)


def test_ddpc_quadrant_differences():
    W, H = 4, 3
    myresul = np.arange(W * H * 4, dtype=np.float64).reshape(4, W * H).T ** 1.5
    px, py, tot = dpc.ddpc(myresul, W, H, norm=2.0)
    img = lambda a: myresul[:, a].reshape(H, W) / 2.0
    np.testing.assert_allclose(px, img(1) - img(3))
    np.testing.assert_allclose(py, img(2) - img(0))
    np.testing.assert_allclose(tot, img(0) + img(1) + img(2) + img(3))


@pytest.mark.gpu
def test_dpc_stem_matches_oracle(engine):
    from oracle import multislice as oms
    from oracle import potential as opot
    from totem_amd import specimen, stem
    print("case dpc rings", flush=True)
    a = 3.905
    x = specimen.perovskite(4, 4, 3, a=a)
    dz = a / 2
    cfg = stem.StemConfig(atoms=x, np_side=256, dx=4 * a / 256, n=128, kv=200.0, alpha_mrad=20.0,
                          slice_thickness=dz, n_slices=6, slice_origin=-dz / 4, paraflag="l",
                          width_red=6, hight_red=4, step=4, begin_row=30, begin_col=40)
    inp = stem.prepare(cfg)
    inp.aper = optics.dpc_masks(cfg.n, cfg.dx, cfg.kv, [(5.0, 20.0, 0.0), (20.0, 45.0, 45.0)])
    inp.aper2 = np.fft.ifftshift(inp.aper, axes=(-2, -1)).copy()
    myresul, _, _, _ = stem.run(inp)
    _, t = opot.potential_slices(cfg.np_side, inp.ele_n, None, inp.ele_n_i, None, inp.series_n, inp.series_n_i,
                                 inp.s2, inp.gx, inp.gy, inp.sigma, inp.parameter, inp.aper_true, "l")
    ref, _ = oms.stem_cpu(t, inp.probe, inp.gfsf, inp.psf, cfg.begin_row, cfg.begin_col, cfg.width_red,
                          cfg.hight_red, cfg.step, inp.aper)
    rel = np.abs(myresul - ref) / np.abs(ref)
    print("  8 quadrant detectors: max rel per pixel %.3e (FP64-oracle error)" % rel.max())
    assert rel.max() <= 1e-4
    for ring in range(2):
        px, py, _ = dpc.ddpc(myresul, cfg.width_red, cfg.hight_red, ring, norm=cfg.n ** 2)
        qx, qy, _ = dpc.ddpc(ref, cfg.width_red, cfg.hight_red, ring, norm=cfg.n ** 2)
        p, q = dpc.idpc(px, py, cfg.dx * 0.1, cfg.step), dpc.idpc(qx, qy, cfg.dx * 0.1, cfg.step)
        err = np.linalg.norm(p - q) / np.linalg.norm(q)
        print("  ring %d: IDPC rel L2 %.3e" % (ring, err))
        assert err <= 1e-4
