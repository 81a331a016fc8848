"""The allocation-free STEM entry point (totem_stem_run_static) against totem_stem_run.

totem_stem_run_static runs a job in the library's static device scratch instead of
an allocated workspace.  The kernels and their order are those of totem_stem_run,
so the detector images must be bit-identical for the same batch; two calls in a
row (the second on a smaller job: smaller grid, fewer slices, probes and
detectors) check that the shared scratch carries nothing from one call into the
next.  A job whose workspace exceeds the scratch is refused before anything runs.
"""
import numpy as np
import pytest

from oracle import multislice as oms
from oracle import potential as opot
from totem_amd import optics, specimen, stem

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _case(flag, n, np_side, S, scan, step, begin, M, mid, dets):
    a = 3.905
    x = specimen.perovskite(4, 4, max(1, S // 2), a=a)
    rng = np.random.default_rng(17)
    x[:, 1:3] += rng.normal(0, 0.07, (len(x), 2))
    dz = a / 2
    ab = optics.Aberrations(C1=-5.0, Cs_um=1.0) if M > 1 else None
    cfg = stem.StemConfig(atoms=x, np_side=np_side, dx=4 * a / np_side, n=n, kv=200.0, alpha_mrad=24.0,
                          slice_thickness=dz, n_slices=S, slice_origin=-dz / 4, paraflag=flag, aberrations=ab,
                          focal_members=M, focal_rms_nm=3.0, detectors_mrad=dets, width_red=scan[0],
                          hight_red=scan[1], step=step, begin_row=begin[0], begin_col=begin[1], mid_slices=mid)
    return stem.prepare(cfg)


def _transmission(inp):
    c = inp.cfg
    kw = {}
    if c.paraflag == "n":
        kw = dict(ele_n_corr=inp.ele_n_corr, corr_matrix=inp.corr_matrix, series_n_corr=inp.series_n_corr)
    _, t = opot.potential_slices(c.np_side, inp.ele_n, None, inp.ele_n_i, None, inp.series_n, inp.series_n_i,
                                 inp.s2, inp.gx, inp.gy, inp.sigma, inp.parameter, inp.aper_true, c.paraflag, **kw)
    return t.astype(np.complex64)


def _job(engine, inp, t_dev):
    torch = _torch()
    c = inp.cfg
    probes = torch.tensor(inp.probe.astype(np.complex64), device="cuda")
    prop = torch.tensor(optics.unit_complex64(inp.psf), device="cuda")
    aper = torch.tensor(inp.aper, device="cuda")
    aper2 = torch.tensor(inp.aper2, device="cuda")
    L, A, M = len(c.mid_slices), inp.aper.shape[0], len(inp.gfsf)
    det = torch.full(((L + 1), A, M * c.depth), -1.0, dtype=torch.float64, device="cuda")
    job = engine.stem_job(c.n, c.np_side, c.n_slices, t_dev.data_ptr(), probes.data_ptr(), M, prop.data_ptr(),
                          aper.data_ptr(), A, c.depth, c.width_red, c.step, c.begin_row - 1, c.begin_col - 1,
                          det.data_ptr(), aper2_ptr=aper2.data_ptr(), mid=sorted(c.mid_slices))
    job._keep = (probes, prop, aper, aper2)
    return job, det


def _check(engine, inp, batch, label):
    torch = _torch()
    c = inp.cfg
    t32 = _transmission(inp)
    t_dev = torch.tensor(t32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    job_a, det_a = _job(engine, inp, t_dev)
    job_s, det_s = _job(engine, inp, t_dev)
    need = engine.stem_workspace_bytes(job_s, batch)
    print("case %s: n=%d np=%d S=%d probes=%d batch=%d workspace %d bytes" %
          (label, c.n, c.np_side, c.n_slices, len(inp.gfsf) * c.depth, batch, need), flush=True)
    assert need <= engine.STATIC_SCRATCH_BYTES
    engine.stem_run(job_a, batch, stream=stream)
    torch.cuda.synchronize()
    engine.stem_run_static(job_s, batch, stream=stream)
    torch.cuda.synchronize()
    a, s = det_a.cpu().numpy(), det_s.cpu().numpy()
    assert np.all(s >= 0.0)
    np.testing.assert_array_equal(s, a)
    # and the exit face agrees with the FP64 oracle fed the same FP32 transmission
    res_o, _ = oms.stem_cpu(t32.astype(np.complex128), inp.probe, inp.gfsf, inp.psf, c.begin_row, c.begin_col,
                            c.width_red, c.hight_red, c.step, inp.aper)
    img = np.einsum("amd,m->da", s[-1].reshape(s.shape[1], len(inp.gfsf), c.depth), inp.gfsf)
    rel = float(np.max(np.abs(img - res_o) / np.abs(res_o)))
    print("  bit-identical to totem_stem_run; max rel err vs FP64 oracle %.2e" % rel, flush=True)
    assert rel <= 1e-4


def test_static_run_equals_allocating_run_twice(engine):
    big = _case("l", n=128, np_side=256, S=6, scan=(4, 3), step=5, begin=(20, 30), M=2, mid=(3,),
                dets=((60.0, 200.0), (0.0, 10.0), (20.0, 40.0)))
    _check(engine, big, batch=5, label="first call")
    small = _case("n", n=64, np_side=128, S=4, scan=(2, 2), step=3, begin=(10, 12), M=1, mid=(),
                  dets=((60.0, 200.0),))
    _check(engine, small, batch=3, label="second call, smaller job")


def test_static_run_refuses_a_job_beyond_the_scratch(engine):
    torch = _torch()
    n = 2048
    t = torch.zeros((1, n, n), dtype=torch.complex64, device="cuda")
    probes = torch.zeros((1, n, n), dtype=torch.complex64, device="cuda")
    prop = torch.zeros((n, n), dtype=torch.complex64, device="cuda")
    aper = torch.zeros((1, n, n), dtype=torch.float64, device="cuda")
    det = torch.zeros((1, 1, 64), dtype=torch.float64, device="cuda")
    job = engine.stem_job(n, n, 1, t.data_ptr(), probes.data_ptr(), 1, prop.data_ptr(), aper.data_ptr(), 1, 64, 8,
                          1, 0, 0, det.data_ptr(), wrap=True)
    assert engine.stem_workspace_bytes(job, 64) > engine.STATIC_SCRATCH_BYTES
    print("case oversized static run: 64 probes of 2048^2", flush=True)
    with pytest.raises(engine.TotemError, match="static scratch") as e:
        engine.stem_run_static(job, 64, stream=torch.cuda.current_stream().cuda_stream)
    assert e.value.code == -3
