"""Precision stress (2048^2, 210 slices): the FP32 engine against an FP64 multislice on the GPU.

Specimen: diamond-cubic Si viewed along [110], ~104 x 104 x 300 A, capped top and bottom by 50 A of
amorphous carbon (2.0 g/cm^3, 1.4 A minimum separation, seeded): 270 000 atoms.  Both paths take the
same transmission stack (built once by the engine, complex64) so the comparison isolates the
propagation arithmetic; the FP64 path is the reference's multislice loop
(``cuda_STEM_CPU_Mcore_lobato_peng.m:60-74``) in complex128 with torch's FFT, the propagator and
probe from the FP64 optics.  The exit wave Psi(k) = FFT2(psi) . H is compared at slice checkpoints:
the per-slice growth of the relative L2 error of Psi and of |Psi|^2 is printed, and the exit-wave
intensity must stay within 1e-5 relative L2 through all 210 slices.
"""
import numpy as np
import pytest

from totem_amd import stem

pytestmark = pytest.mark.gpu

CHECKPOINTS = (1, 2, 5, 10, 20, 50, 100, 150, 210)
BEGIN = (700, 1300)
STEP = 37


def _fp64_exit_waves(torch, t, probe, psf, r0, c0, checkpoints):
    psi = torch.tensor(probe, dtype=torch.complex128, device="cuda")
    H = torch.tensor(psf, dtype=torch.complex128, device="cuda")
    out = {}
    for s in range(max(checkpoints)):
        psi = psi * torch.roll(t[s], shifts=(-r0, -c0), dims=(0, 1)).to(torch.complex128)
        Psi = torch.fft.fft2(psi) * H
        if s + 1 in checkpoints:
            out[s + 1] = Psi.clone()
        psi = torch.fft.ifft2(Psi)
    return out


def test_precision_stress_fp32_against_fp64(engine):
    import torch
    from totem_amd import optics
    assert torch.cuda.is_available()
    print("case precision stress 2048^2 x 210 slices", flush=True)
    cfg = stem.precision_c3()
    inp = stem.prepare(cfg)
    S, N = cfg.n_slices, cfg.n
    assert S == 210 and N == 2048 and len(cfg.atoms) > 250000
    st = torch.cuda.current_stream().cuda_stream
    t = torch.empty((S, N, N), dtype=torch.complex64, device="cuda")
    engine.build_transmission("l", N, S, inp.series_n, inp.ele_n, inp.gx, inp.gy, inp.s2, inp.aper_true, inp.sigma,
                              inp.parameter, t.data_ptr(), stream=st)
    probes = torch.tensor(inp.probe.astype(np.complex64), device="cuda")
    prop = torch.tensor(optics.unit_complex64(inp.psf), device="cuda")
    aper = torch.tensor(inp.aper, device="cuda")
    exit_wave = torch.zeros((2, N, N), dtype=torch.complex64, device="cuda")
    det = torch.zeros((1, 1, 2), dtype=torch.float64, device="cuda")
    fp32 = {}
    for k in CHECKPOINTS:
        job = engine.stem_job(N, N, k, t.data_ptr(), probes.data_ptr(), 1, prop.data_ptr(), aper.data_ptr(), 1, 2, 2,
                              STEP, BEGIN[0], BEGIN[1], det.data_ptr(), wrap=True, exit_ptr=exit_wave.data_ptr())
        engine.stem_run(job, 2, stream=st)
        torch.cuda.synchronize()
        fp32[k] = exit_wave.clone()
    worst = 0.0
    for d in range(2):
        ref = _fp64_exit_waves(torch, t, inp.probe[0], inp.psf, BEGIN[0], BEGIN[1] + d * STEP, CHECKPOINTS)
        print("  position %d: slice, rel L2 of Psi, rel L2 of |Psi|^2, |Psi|^2 sum / N^2" % d)
        for k in CHECKPOINTS:
            a = fp32[k][d].to(torch.complex128)
            b = ref[k]
            e_psi = float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
            ia, ib = a.abs() ** 2, b.abs() ** 2
            e_int = float(torch.linalg.vector_norm(ia - ib) / torch.linalg.vector_norm(ib))
            print("    %4d  %.3e  %.3e  %.8f" % (k, e_psi, e_int, float(ib.sum()) / N**2))
            worst = max(worst, e_int)
        del ref
    assert worst <= 1e-5
