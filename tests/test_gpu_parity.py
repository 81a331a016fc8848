"""GPU parity: the HIP engine through its C ABI against the FP64 oracle on identical inputs.

Tolerances (stated by the engine's contract): potential slices <= 1e-6 relative L2,
transmission <= 1e-6 relative L2, annular-detector images <= 1e-4 relative per pixel;
batched / plain / graph-replayed runs are bit-identical.  FP64-oracle errors are
printed for every case.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import multislice as oms
from oracle import potential as opot
from totem_amd import optics, specimen, stem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _stream():
    return _torch().cuda.current_stream().cuda_stream


def _rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


# ---------------------------------------------------------------------------- FFT

@pytest.mark.parametrize("n", [64, 128, 256, 320, 512, 768, 1024, 1536, 2048])
def test_fft2d_matches_numpy(engine, n):
    torch = _torch()
    print("case fft2d n=%d" % n, flush=True)
    rng = np.random.default_rng(n)
    batch = 2 if n <= 1024 else 1
    x = (rng.standard_normal((batch, n, n)) + 1j * rng.standard_normal((batch, n, n))).astype(np.complex64)
    for sign in (-1, +1):
        d = torch.tensor(x, device="cuda")
        engine.fft2d(d.data_ptr(), n, batch, sign, stream=_stream())
        torch.cuda.synchronize()
        y = d.cpu().numpy().astype(np.complex128)
        ref = np.fft.fft2(x.astype(np.complex128)) if sign < 0 else np.fft.ifft2(x.astype(np.complex128)) * n * n
        err = max(_rel_l2(y[b], ref[b]) for b in range(batch))
        print("  sign %+d rel L2 %.3e" % (sign, err))
        assert err < 2e-6


# ---------------------------------------------------------------------- potential

def _displaced_perovskite(nx, ny, nz, seed, occ=True):
    a = 3.905
    x = specimen.perovskite(nx, ny, nz, a=a)
    rng = np.random.default_rng(seed)
    x[:, 1:3] += rng.normal(0, 0.07, (len(x), 2))
    if occ:
        x[::7, 7] = 0.8
    return x


def _potential_case(flag, np_side=128, nx=2, S=4, ions=False, absorption=False, seed=1):
    a = 3.905
    x = _displaced_perovskite(nx, nx, S // 2, seed)
    if ions:
        o = np.nonzero(x[:, 0] == 8)[0][::3]
        x[o, 4] = -2
        x[o, 5] = 0.5
    dz = a / 2
    cfg = stem.StemConfig(atoms=x, np_side=np_side, dx=nx * a / np_side, n=64, kv=300.0, alpha_mrad=24.0,
                          slice_thickness=dz, n_slices=S, slice_origin=-dz / 4, paraflag=flag)
    inp = stem.prepare(cfg)
    absorp = None
    if absorption:
        absorp = np.tile(np.array([0.01, 2.0, 0.02, 5.0, 0.005, 20.0, 0.0, 1.0, 0.0, 1.0]), (len(inp.ele_n), 1))
    return cfg, inp, absorp


def _build_gpu(engine, inp, absorp=None):
    torch = _torch()
    c = inp.cfg
    S, npx = c.n_slices, c.np_side
    t = torch.empty((S, npx, npx), dtype=torch.complex64, device="cuda")
    corr = c.paraflag == "n"
    vre, vim = engine.build_transmission(
        c.paraflag, npx, S, inp.series_n, inp.ele_n, inp.gx, inp.gy, inp.s2, inp.aper_true, inp.sigma,
        inp.parameter, t.data_ptr(), absorp_n=absorp,
        atom_slice_i=inp.series_n_i if len(inp.series_n_i) else None,
        proj_coff_mat_i=inp.ele_n_i if len(inp.ele_n_i) else None,
        series_n_corr=inp.series_n_corr if corr else None, ele_n_corr=inp.ele_n_corr if corr else None,
        corr_info_matrix=inp.corr_matrix if corr else None, want_v=True, stream=_stream())
    torch.cuda.synchronize()
    return vre + 1j * vim, t.cpu().numpy().astype(np.complex128)


def _oracle_potential(inp, absorp=None):
    c = inp.cfg
    kw = {}
    if c.paraflag == "n":
        kw = dict(ele_n_corr=inp.ele_n_corr, corr_matrix=inp.corr_matrix, series_n_corr=inp.series_n_corr)
    return opot.potential_slices(c.np_side, inp.ele_n, absorp, inp.ele_n_i, None, inp.series_n, inp.series_n_i,
                                 inp.s2, inp.gx, inp.gy, inp.sigma, inp.parameter, inp.aper_true, c.paraflag, **kw)


@pytest.mark.parametrize("flag,ions,absorption", [("l", False, False), ("p", False, False), ("n", False, False),
                                                  ("p", True, True), ("l", False, True)])
def test_potential_matches_oracle(engine, flag, ions, absorption):
    print("case potential flag=%s ions=%s absorption=%s" % (flag, ions, absorption), flush=True)
    cfg, inp, absorp = _potential_case(flag, ions=ions, absorption=absorption)
    if ions:
        assert len(inp.ele_n_i) > 0
    V, t = _build_gpu(engine, inp, absorp)
    Vo, to = _oracle_potential(inp, absorp)
    for s in range(cfg.n_slices):
        ev, et = _rel_l2(V[s], Vo[s]), _rel_l2(t[s], to[s])
        print("  slice %d: V rel L2 %.3e  t rel L2 %.3e" % (s, ev, et))
        assert ev <= 1e-6
        assert et <= 1e-6


@pytest.mark.parametrize("flag,np_side,shift", [("p", 320, 0.0), ("n", 256, 0.0), ("p", 128, 0.25),
                                                 ("n", 128, 0.25)])
def test_potential_paired_and_unpaired_grids_match_oracle(engine, flag, np_side, shift):
    """Paired frequencies (k[np - i] = -k[i]): a quarter of the tiles is accumulated and each
    one's mirrored pixels are written from the same products; per-pixel s2 and AperTrue are read
    at every pixel, so both are made asymmetric here.  A grid shifted by a fraction of a pixel has
    no pairs and takes the full accumulation.  Both within 1e-6 of the oracle."""
    from totem_amd import engine as eng
    print("case potential flag=%s np=%d shift=%.2f" % (flag, np_side, shift), flush=True)
    cfg, inp, _ = _potential_case(flag, np_side=np_side)
    rng = np.random.default_rng(np_side)
    dk = inp.gx[0, 1] - inp.gx[0, 0]
    inp.gx = inp.gx + shift * dk
    inp.s2 = inp.s2 * (1.0 + 0.05 * rng.random(inp.s2.shape))
    inp.aper_true = (rng.random(inp.s2.shape) > 0.2).astype(np.float64)
    c = inp.cfg
    st = eng.potential_plan_stats(c.paraflag, c.np_side, c.n_slices, inp.series_n, inp.ele_n, inp.gx, inp.gy,
                                  inp.s2, inp.aper_true, inp.sigma, inp.parameter)
    assert st["paired"] == (shift == 0.0)
    V, t = _build_gpu(engine, inp)
    Vo, to = _oracle_potential(inp)
    for s in range(cfg.n_slices):
        ev, et = _rel_l2(V[s], Vo[s]), _rel_l2(t[s], to[s])
        print("  slice %d: V rel L2 %.3e  t rel L2 %.3e" % (s, ev, et))
        assert ev <= 1e-6
        assert et <= 1e-6


def test_potential_with_per_atom_coefficients_matches_oracle(engine):
    """Peng rows that all differ (per-atom Debye-Waller B folded into b, per-atom absorptive
    coefficients, mixed ions): the planner splits them into per-Gaussian classes."""
    print("case potential per-atom coefficients (Gaussian split)", flush=True)
    from dataclasses import replace
    cfg, inp, _ = _potential_case("p", ions=True)
    rng = np.random.default_rng(8)
    x = cfg.atoms.copy()
    x[:, 6] = rng.uniform(0.2, 0.8, len(x))
    cfg = replace(cfg, atoms=x, dw_selected=True)
    inp = stem.prepare(cfg)
    absorp = np.abs(rng.normal(0.0, 0.02, (len(inp.ele_n), 10))) * np.tile([1.0, 200.0], 5)
    assert len({tuple(r[3:]) for r in inp.ele_n}) > len(inp.ele_n) // 2
    V, t = _build_gpu(engine, inp, absorp)
    Vo, to = _oracle_potential(inp, absorp)
    for s in range(cfg.n_slices):
        ev, et = _rel_l2(V[s], Vo[s]), _rel_l2(t[s], to[s])
        print("  slice %d: V rel L2 %.3e  t rel L2 %.3e" % (s, ev, et))
        assert ev <= 1e-6
        assert et <= 1e-6


def test_potential_only_entry_point_matches_oracle(engine):
    print("case waveCudaPengAndCorrP", flush=True)
    cfg, inp, _ = _potential_case("n")
    cfg.width_red = cfg.hight_red = 1
    _, _, px, py = stem.run(inp, potential_only=True)
    _, to = _oracle_potential(inp)
    S, npx = cfg.n_slices, cfg.np_side
    t = (px + 1j * py).T.reshape(S, npx, npx)
    err = _rel_l2(t, to)
    print("  t rel L2 %.3e" % err)
    assert err <= 1e-6


# --------------------------------------------------------------------------- STEM

def _stem_case(flag, M=1, mid=(), n=128, np_side=256, S=6, cells=4, scan=(3, 2), step=5, begin=(20, 30), seed=4):
    a = 3.905
    x = _displaced_perovskite(cells, cells, S // 2, seed)
    dz = a / 2
    ab = optics.Aberrations(C1=-5.0, Cs_um=1.0) if M > 1 else None
    cfg = stem.StemConfig(atoms=x, np_side=np_side, dx=cells * a / np_side, n=n, kv=300.0 if flag != "l" else 200.0,
                          alpha_mrad=24.0, slice_thickness=dz, n_slices=S, slice_origin=-dz / 4, paraflag=flag,
                          aberrations=ab, focal_members=M, focal_rms_nm=3.0,
                          detectors_mrad=((60.0, 200.0), (0.0, 10.0), (20.0, 40.0)), width_red=scan[0],
                          hight_red=scan[1], step=step, begin_row=begin[0], begin_col=begin[1], mid_slices=mid)
    return cfg, stem.prepare(cfg)


def _oracle_stem(inp, t=None):
    c = inp.cfg
    if t is None:
        _, t = _oracle_potential(inp)
    return oms.stem_cpu(t, inp.probe, inp.gfsf, inp.psf, c.begin_row, c.begin_col, c.width_red, c.hight_red,
                        c.step, inp.aper, tuple(sorted(c.mid_slices)))


def _check_images(res, ref, what):
    rel = np.abs(res - ref) / np.abs(ref)
    print("  %s: max rel per pixel %.3e (FP64-oracle error), min value %.3e" % (what, rel.max(), np.abs(ref).min()))
    assert rel.max() <= 1e-4


@pytest.mark.parametrize("flag,M,mid", [("l", 1, ()), ("n", 3, (2, 5)), ("p", 1, (3,))])
def test_stem_myprojectfun_matches_oracle(engine, flag, M, mid):
    print("case stem flag=%s M=%d mid=%s" % (flag, M, mid), flush=True)
    cfg, inp = _stem_case(flag, M=M, mid=mid)
    myresul, mid_mat, px, py = stem.run(inp)
    res_o, mid_o = _oracle_stem(inp)
    assert myresul.shape == res_o.shape and mid_mat.shape == mid_o.shape
    _check_images(myresul, res_o, "exit-face detectors")
    if mid:
        _check_images(mid_mat, mid_o, "mid-slice detectors")
    assert np.all(px == 0) and np.all(py == 0)


@pytest.mark.parametrize("flag,M,mid", [("l", 1, (2,)), ("n", 3, (2, 5))])
def test_drop_in_library_symbol_matches_oracle(engine, flag, M, mid):
    """``myProjectfun`` exported by libwaveCudalobato.so / libwaveCudaPengAndCorr.so (the symbol the
    mex gateway binds, waveCuda.h:10-21) called through ctypes: the same images as the engine entry
    point bit for bit, within 1e-4 per pixel of the FP64 oracle."""
    print("case drop-in symbol %s flag=%s M=%d mid=%s" % (stem.mex_name(flag), flag, M, mid), flush=True)
    cfg, inp = _stem_case(flag, M=M, mid=mid)
    res_d, mid_d, px, py = stem.run(inp, drop_in=True)
    res_e, mid_e, _, _ = stem.run(inp)
    np.testing.assert_array_equal(res_d, res_e)
    np.testing.assert_array_equal(mid_d, mid_e)
    res_o, mid_o = _oracle_stem(inp)
    _check_images(res_d, res_o, "exit-face detectors")
    _check_images(mid_d, mid_o, "mid-slice detectors")
    assert np.all(px == 0) and np.all(py == 0)


@pytest.mark.parametrize("flag", ["l", "n"])
def test_drop_in_potential_library_symbol_matches_oracle(engine, flag):
    """``myProjectfun`` of libwaveCudalobatoP.so / libwaveCudaPengAndCorrP.so: the transmission
    function within 1e-6 relative L2 of the oracle, equal to the engine entry point's."""
    print("case drop-in symbol %s" % stem.mex_name(flag, True), flush=True)
    cfg, inp, _ = _potential_case(flag)
    cfg.width_red = cfg.hight_red = 1
    _, _, px, py = stem.run(inp, potential_only=True, drop_in=True)
    _, _, pxe, pye = stem.run(inp, potential_only=True)
    np.testing.assert_array_equal(px, pxe)
    np.testing.assert_array_equal(py, pye)
    _, to = _oracle_potential(inp)
    S, npx = cfg.n_slices, cfg.np_side
    err = _rel_l2((px + 1j * py).T.reshape(S, npx, npx), to)
    print("  t rel L2 %.3e" % err)
    assert err <= 1e-6


def test_drop_in_library_symbol_reports_errors(engine):
    """A window that leaves the potential: the drop-in symbol returns nothing (as the reference's
    does) and leaves the message in totem_last_error(), which the wrapper raises."""
    cfg, inp = _stem_case("l", scan=(3, 2), step=60, begin=(20, 30))
    with pytest.raises(engine.TotemError, match="window leaves the potential"):
        stem.run(inp, drop_in=True)


@pytest.mark.parametrize("flag,n,np_side,S", [("n", 320, 512, 4), ("n", 768, 1024, 4), ("l", 1536, 2048, 2)])
def test_stem_mixed_radix_grids_match_oracle(engine, flag, n, np_side, S):
    """Windows of 320 (radix 5), 768 and 1536 (radix 3) points -- the plans whose rows pass exchanges
    with wave-local barriers -- inside a larger potential, with a mid-slice output."""
    print("case stem mixed-radix n=%d np=%d S=%d flag=%s" % (n, np_side, S, flag), flush=True)
    cfg, inp = _stem_case(flag, M=1, mid=(1,), n=n, np_side=np_side, S=S, cells=2, scan=(2, 1), step=7,
                          begin=(11, 23))
    myresul, mid_mat, _, _ = stem.run(inp)
    res_o, mid_o = _oracle_stem(inp)
    _check_images(myresul, res_o, "exit-face detectors")
    _check_images(mid_mat, mid_o, "mid-slice detectors")


@pytest.mark.parametrize("n,shift", [(128, 64), (128, 37), (768, 300)])
def test_stem_with_shifted_propagator_band_matches_oracle(engine, n, shift):
    """A propagator whose nonzero columns form a band that does (shift 37) or does not (shift
    n/2) wrap through column 0: the passes process that band only (the rest of the wave is zero
    after every propagator multiply) and the images match the oracle run with the same propagator."""
    import dataclasses
    print("case shifted band n=%d shift=%d" % (n, shift), flush=True)
    cfg, inp = _stem_case("l", M=1, mid=(2,), n=n, np_side=256 if n <= 128 else 1024, cells=2 if n > 128 else 4,
                          scan=(2, 1), step=7, begin=(11, 23))
    inp = dataclasses.replace(inp, psf=np.roll(inp.psf, shift, axis=1))
    lo, ln = engine.stem_band(inp.psf, engine.plan_columns(n, engine.plan(n)[0]))
    assert 0 < ln < n
    myresul, mid_mat, _, _ = stem.run(inp)
    res_o, mid_o = _oracle_stem(inp)
    # a shifted band can leave a detector without signal (the oracle's FP64 roundoff there is
    # ~1e-35): relative error per pixel with a floor of 1e-6 of the image's peak
    for res, ref, what in ((myresul, res_o, "exit-face detectors"), (mid_mat, mid_o, "mid-slice detectors")):
        rel = np.abs(res - ref) / (np.abs(ref) + 1e-6 * np.abs(ref).max())
        print("  %s: max rel per pixel %.3e" % (what, rel.max()))
        assert rel.max() <= 1e-4


def test_band_limited_run_equals_full_grid_run(engine):
    """The passes skip the columns where the propagator is zero: the same job with 1e-30 in
    every zero entry of the propagator (so no column is zero and the whole grid is transformed)
    gives the same detector images to 1e-6 -- a check of the skipping that does not go through
    the oracle's tolerance."""
    import dataclasses
    print("case band vs full grid", flush=True)
    cfg, inp = _stem_case("n", M=1, mid=(2, 4), n=512, np_side=1024, cells=2, scan=(2, 2), step=9,
                          begin=(11, 23))
    C = engine.plan_columns(512, engine.plan(512)[0])
    full = np.where(inp.psf == 0, 1e-30, inp.psf)
    assert engine.stem_band(inp.psf, C)[1] < 512 and engine.stem_band(full, C)[1] == 512
    r_band, m_band, _, _ = stem.run(inp)
    r_full, m_full, _, _ = stem.run(dataclasses.replace(inp, psf=full))
    np.testing.assert_allclose(r_band, r_full, rtol=1e-6, atol=0)
    np.testing.assert_allclose(m_band, m_full, rtol=1e-6, atol=0)


def test_zero_propagator_gives_zero_images(engine):
    """A propagator without a nonzero entry: an empty band, every detector image exactly zero (as
    the oracle's)."""
    import dataclasses
    print("case zero propagator", flush=True)
    cfg, inp = _stem_case("l", M=1, mid=(2,), scan=(2, 2), step=5)
    inp = dataclasses.replace(inp, psf=np.zeros_like(inp.psf))
    assert engine.stem_band(inp.psf, 4) == (0, 0)
    myresul, mid_mat, _, _ = stem.run(inp)
    res_o, mid_o = _oracle_stem(inp)
    assert np.all(res_o == 0) and np.all(mid_o == 0)
    assert np.all(myresul == 0) and np.all(mid_mat == 0)


def test_empty_slices_are_vacuum(engine):
    """Slices without atoms (here the last four of six) transmit exactly: t = 1 + 0i, and the scan
    images still match the oracle."""
    print("case empty slices", flush=True)
    cfg, inp = _stem_case("l", S=6, scan=(2, 2), step=5)
    keep = cfg.atoms[:, 3] < cfg.slice_origin + 2 * cfg.slice_thickness
    cfg.atoms = cfg.atoms[keep]
    inp = stem.prepare(cfg)
    assert np.count_nonzero(inp.series_n[2:]) == 0 and np.count_nonzero(inp.series_n[:2]) == 2
    _, t = _build_gpu(engine, inp)
    assert np.all(t[2:] == 1.0)
    myresul, _, _, _ = stem.run(inp)
    res_o, _ = _oracle_stem(inp)
    _check_images(myresul, res_o, "exit-face detectors")


def test_stem_window_outside_potential_is_an_error(engine):
    cfg, inp = _stem_case("l", scan=(3, 2), step=60, begin=(20, 30))
    with pytest.raises(engine.TotemError, match="window leaves the potential"):
        stem.run(inp)


def _device_job(engine, inp, t_dev, wrap=False, begin=None, exit_wave=None):
    torch = _torch()
    c = inp.cfg
    probes = torch.tensor(inp.probe.astype(np.complex64), device="cuda")
    prop = torch.tensor(optics.unit_complex64(inp.psf), device="cuda")
    aper = torch.tensor(inp.aper, device="cuda")
    aper2 = torch.tensor(inp.aper2, device="cuda")
    L, A, M = len(c.mid_slices), inp.aper.shape[0], len(inp.gfsf)
    det = torch.zeros(((L + 1), A, M * c.depth), dtype=torch.float64, device="cuda")
    br, bc = (c.begin_row - 1, c.begin_col - 1) if begin is None else begin
    job = engine.stem_job(c.n, c.np_side, c.n_slices, t_dev.data_ptr(), probes.data_ptr(), M, prop.data_ptr(),
                          aper.data_ptr(), A, c.depth, c.width_red, c.step, br, bc, det.data_ptr(),
                          aper2_ptr=aper2.data_ptr(), mid=sorted(c.mid_slices), wrap=wrap,
                          exit_ptr=None if exit_wave is None else exit_wave.data_ptr())
    job._keep = (probes, prop, aper, aper2)
    return job, det


def test_batched_plain_and_graph_runs_are_bit_identical(engine):
    torch = _torch()
    print("case batch/plain/graph", flush=True)
    cfg, inp = _stem_case("l", M=1, mid=(3,), scan=(4, 3), step=4)
    _, to = _oracle_potential(inp)
    t_dev = torch.tensor(to.astype(np.complex64), device="cuda")
    outs = []
    for batch, graph in ((cfg.depth, False), (1, False), (5, False), (cfg.depth, True), (5, True)):
        job, det = _device_job(engine, inp, t_dev)
        engine.stem_run(job, batch, graph=graph, stream=_stream())
        torch.cuda.synchronize()
        outs.append(det.cpu().numpy())
    for o in outs[1:]:
        np.testing.assert_array_equal(o, outs[0])
    # and the device path agrees with the oracle fed the same (FP32-rounded) transmission
    res_o, mid_o = _oracle_stem(inp, t=to.astype(np.complex64).astype(np.complex128))
    _check_images(outs[0][1].T, res_o, "device API exit face")


def _plan_runs(engine, n, inp, t_dev, plans):
    """Exit-face detector images [A, M*D] of the job run with each listed pass plan."""
    torch = _torch()
    prev, _ = engine.plan(n)
    outs = []
    try:
        for p in plans:
            engine.set_plan(n, p)
            job, det = _device_job(engine, inp, t_dev, wrap=True, begin=(0, 0))
            engine.stem_run(job, inp.cfg.depth, stream=_stream())
            torch.cuda.synchronize()
            outs.append(det.cpu().numpy()[-1])
    finally:
        engine.set_plan(n, prev)
    return outs


@pytest.mark.parametrize("n", [256, 512, 768, 1024, 1536, 2048])
def test_every_pass_plan_matches_oracle(engine, n):
    """Every compiled schedule of the two passes (radix plans, rows / columns per workgroup,
    exchange form, when the rows pass loads the transmission) gives the same images within FP32
    rounding, each within 1e-4 of the FP64 C++ oracle (potential and propagation)."""
    torch = _torch()
    count = engine.plan_count(n)
    print("case pass plans n=%d: %d plans" % (n, count), flush=True)
    assert count >= 2
    cfg, inp = _stem_case("l", n=n, np_side=n, S=4, cells=4, scan=(2, 2), step=7, begin=(1, 1))
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    from oracle import cpu_stem
    threads = max(1, min(16, os.cpu_count() or 1))
    _, to = cpu_stem.potential_slices(n, inp.ele_n, None, None, None, inp.series_n, None, inp.s2, inp.gx, inp.gy,
                                      inp.sigma, inp.parameter, inp.aper_true, cfg.paraflag, threads=threads,
                                      want_v=False)
    t32 = to.astype(np.complex64)
    t_dev = torch.tensor(t32, device="cuda")
    outs = _plan_runs(engine, n, inp, t_dev, range(count))
    ash = np.fft.ifftshift(inp.aper, axes=(-2, -1))
    r0 = [(d // cfg.width_red) * cfg.step for d in range(cfg.depth)]
    c0 = [(d % cfg.width_red) * cfg.step for d in range(cfg.depth)]
    ref, _ = cpu_stem.stem_positions(t32, inp.probe, inp.gfsf, inp.psf, ash, r0, c0, periodic=True,
                                     threads=threads)
    for p, o in enumerate(outs):
        got = o.T
        rel = float(np.max(np.abs(got - ref) / np.abs(ref)))
        same = float(np.max(np.abs(o - outs[0]) / np.abs(outs[0])))
        print("  plan %d: vs FP64 oracle %.2e, vs plan 0 %.2e" % (p, rel, same))
        assert rel <= 1e-4
        assert same <= 1e-5


def test_autotuned_plan_is_adopted(engine):
    """totem_stem_autotune measures plan x batch; the chosen pair is then what runs use."""
    torch = _torch()
    n = 768
    prev, _ = engine.plan(n)
    try:
        batch, rate = engine.autotune(n, slices=4)
        plan, desc = engine.plan(n)
        print("case autotune n=%d: plan %d (%s), batch %d, %.0f probe.slice/s" % (n, plan, desc, batch, rate),
              flush=True)
        assert 0 <= plan < engine.plan_count(n) and desc
        assert engine.preferred_batch(n) == batch and rate > 0
        cfg, inp = _stem_case("l", n=n, np_side=n, S=3, cells=4, scan=(2, 2), step=7, begin=(1, 1))
        _, to = _oracle_potential(inp)
        t_dev = torch.tensor(to.astype(np.complex64), device="cuda")
        job, det = _device_job(engine, inp, t_dev, wrap=True, begin=(0, 0))
        engine.stem_run(job, cfg.depth, stream=_stream())         # the adopted plan, not set here
        torch.cuda.synchronize()
        adopted = det.cpu().numpy()[-1]
        explicit = _plan_runs(engine, n, inp, t_dev, [plan])[0]
        np.testing.assert_array_equal(adopted, explicit)
    finally:
        engine.set_plan(n, prev)


def _exit_wave_errors(inp, t_host, ew, det, begin, wrap):
    """Exit-wave intensity rel L2 of every (member, position) and the detector errors against
    the oracle fed the same FP32 transmission."""
    c = inp.cfg
    M, D = len(inp.gfsf), c.depth
    ash = np.fft.ifftshift(inp.aper, axes=(-2, -1))
    worst = 0.0
    ref = np.zeros((D, inp.aper.shape[0]))
    for d in range(D):
        r0 = begin[0] + (d // c.width_red) * c.step
        c0 = begin[1] + (d % c.width_red) * c.step
        r, _, waves = oms.stem_position(t_host, inp.probe, inp.gfsf, inp.psf, ash, r0, c0, periodic=wrap, waves=True)
        ref[d] = r
        for m in range(M):
            got = np.abs(ew[m * D + d].astype(np.complex128)) ** 2
            exp = np.abs(waves[m]) ** 2
            worst = max(worst, float(np.linalg.norm(got - exp) / np.linalg.norm(exp)))
    return worst, ref


def test_exit_wave_intensity_matches_oracle(engine):
    torch = _torch()
    print("case exit wave 256^2 x 24 slices, Peng+corr, M=2", flush=True)
    cfg, inp = _stem_case("n", M=2, n=256, np_side=256, S=24, cells=4, scan=(2, 2), step=9, begin=(1, 1))
    _, to = _oracle_potential(inp)
    t32 = to.astype(np.complex64)
    t_dev = torch.tensor(t32, device="cuda")
    ew = torch.zeros((len(inp.gfsf) * cfg.depth, cfg.n, cfg.n), dtype=torch.complex64, device="cuda")
    job, det = _device_job(engine, inp, t_dev, wrap=True, begin=(0, 0), exit_wave=ew)
    engine.stem_run(job, 3, stream=_stream())
    torch.cuda.synchronize()
    worst, _ = _exit_wave_errors(inp, t32.astype(np.complex128), ew.cpu().numpy(), det, (0, 0), True)
    print("  exit-wave intensity: max rel L2 %.3e (FP64-oracle error)" % worst)
    assert worst <= 1e-5


def test_exit_wave_full_size_phonon_configuration(engine):
    """1024^2, 80 slices of a frozen-phonon SrTiO3 configuration (80 000 atoms, Peng + correction):
    the exit wave of one probe against the FP64 oracle on the same transmission."""
    torch = _torch()
    print("case exit wave 1024^2 x 80 slices (phonon configuration)", flush=True)
    cfg = stem.perovskite_c2(scan=1)
    snap = stem.phonon_snapshots(cfg, 1, stem.C2_PHONON_SIGMA, seed=1002)[0]
    rows = stem.prepare_specimen(cfg, snap)
    inp = stem.prepare(cfg)
    S, N = cfg.n_slices, cfg.n
    t = torch.empty((S, N, N), dtype=torch.complex64, device="cuda")
    engine.build_transmission("n", N, S, rows["series_n"], rows["ele_n"], inp.gx, inp.gy, inp.s2, inp.aper_true,
                              inp.sigma, inp.parameter, t.data_ptr(), series_n_corr=rows["series_n_corr"],
                              ele_n_corr=rows["ele_n_corr"], corr_info_matrix=rows["corr_matrix"], stream=_stream())
    ew = torch.zeros((1, N, N), dtype=torch.complex64, device="cuda")
    job, det = _device_job(engine, inp, t, wrap=True, begin=(500, 300), exit_wave=ew)
    engine.stem_run(job, 1, stream=_stream())
    torch.cuda.synchronize()
    t_host = t.cpu().numpy().astype(np.complex128)
    worst, ref = _exit_wave_errors(inp, t_host, ew.cpu().numpy(), det, (500, 300), True)
    del t_host
    print("  exit-wave intensity: rel L2 %.3e (FP64-oracle error)" % worst)
    assert worst <= 1e-5
    _check_images(det.cpu().numpy()[0].T, ref, "ADF/BF")


def test_periodic_windows_equal_tiled_potential(engine):
    torch = _torch()
    print("case periodic windows", flush=True)
    cfg, inp = _stem_case("l", n=128, np_side=128, cells=2, S=4, scan=(3, 3), step=40, begin=(1, 1))
    _, to = _oracle_potential(inp)
    t_dev = torch.tensor(to.astype(np.complex64), device="cuda")
    job, det = _device_job(engine, inp, t_dev, wrap=True, begin=(100, 90))
    engine.stem_run(job, cfg.depth, stream=_stream())
    torch.cuda.synchronize()
    tiled = np.tile(to.astype(np.complex64).astype(np.complex128), (1, 2, 2))
    ref = np.zeros((cfg.depth, inp.aper.shape[0]))
    for d in range(cfg.depth):
        r0 = (100 + (d // cfg.width_red) * cfg.step) % 128
        c0 = (90 + (d % cfg.width_red) * cfg.step) % 128
        r, _ = oms.stem_position(tiled, inp.probe, inp.gfsf, inp.psf, np.fft.ifftshift(inp.aper, axes=(-2, -1)),
                                 r0, c0)
        ref[d] = r
    _check_images(det.cpu().numpy()[0].T, ref, "periodic windows")


def test_full_size_vacuum_closed_form(engine):
    """1024^2, 80 slices, periodic windows: for t = 1 the exit-face detectors equal
    sum_k |FFT(probe)|^2 over (band limit x mask), and a full-band BF collects N^2."""
    torch = _torch()
    print("case vacuum 1024^2 x 80 slices", flush=True)
    n, S, dx, kv = 1024, 80, 78.1 / 1024, 300.0
    probe, gfsf = optics.probe_stack(n, dx, kv, 24.0, optics.Aberrations(C1=-5.0, Cs_um=1.0))
    psf = optics.propagator(n, dx, kv, 3.905 / 2)
    aper = optics.detector_masks(n, dx, kv, [(10.0, 20.0), (0.0, 10.0), (0.0, 500.0)])
    t = torch.ones((S, n, n), dtype=torch.complex64, device="cuda")
    cfg = stem.StemConfig(atoms=np.zeros((0, 8)), np_side=n, dx=dx, n=n, kv=kv, alpha_mrad=24, slice_thickness=1.95,
                          n_slices=S, width_red=2, hight_red=2, step=3)
    none = {k: None for k in ("ele_n", "ele_n_i", "series_n", "series_n_i", "ele_n_corr", "series_n_corr",
                              "corr_info", "corr_matrix", "gx", "gy", "s2", "aper_true", "parameter", "sigma")}
    inp = stem.StemInputs(cfg=cfg, probe=probe, gfsf=gfsf, psf=psf, aper=aper,
                          aper2=np.fft.ifftshift(aper, axes=(-2, -1)), **none)
    job, det = _device_job(engine, inp, t, wrap=True, begin=(0, 0))
    engine.stem_run(job, 4, stream=_stream())
    torch.cuda.synchronize()
    out = det.cpu().numpy()[0]
    P = np.fft.fftshift(np.fft.fft2(probe[0]))
    band = np.fft.fftshift(np.abs(psf) > 0)
    closed = np.array([np.sum(np.abs(P) ** 2 * band * aper[a]) for a in range(3)])
    for d in range(4):
        rel = np.abs(out[:, d] - closed) / closed
        print("  position %d rel err %s" % (d, rel))
        assert rel.max() < 1e-4
    print("  full-band intensity / N^2 - 1 = %.3e" % (out[2, 0] / n ** 2 - 1))
    assert out[2, 0] / n ** 2 == pytest.approx(1.0, abs=1e-5)


def test_autotune_adopts_the_measured_batch(engine):
    print("case autotune 256^2", flush=True)
    rule = engine.preferred_batch(256)
    b, rate = engine.autotune(256, 4)
    print("  cache-fit rule %d, measured best %d (%.0f probe.slice/s)" % (rule, b, rate))
    assert rule // 2 <= b <= 2 * rule and rate > 0
    assert engine.preferred_batch(256) == b


def test_small_lobato_workload_full_scan_matches_oracle(engine):
    """The small Lobato workload (SrTiO3 5x5x10 cells, 1250 atoms, 256^2 over 19.525 A, 20 slices,
    200 kV, 20 mrad, ADF 60-200 mrad, 16 x 16 positions at 3 px) with 128^2 probe windows inside the
    potential (no window wraps), through myProjectfun, against the FP64 oracle per pixel."""
    print("case small Lobato workload 16x16 scan", flush=True)
    cfg = stem.perovskite_c1(n=128, np_side=256, scan=16, step=3)
    cfg.begin_row, cfg.begin_col = 41, 37
    inp = stem.prepare(cfg)
    myresul, _, _, _ = stem.run(inp)
    ref, _ = _oracle_stem(inp)
    assert myresul.shape == (256, 1)
    _check_images(myresul, ref, "ADF 60-200 mrad image (256 positions)")
