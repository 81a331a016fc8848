"""The FP64 C++ oracle (oracle/cpu_stem.cpp, the CPU baseline of bench.py) against the
numpy oracle it restates, and its FFT against numpy's.  CPU only.

Both follow GetPotential4AllSlice_multicore_lotabo_peng(_corr).m and
cuda_STEM_CPU_Mcore_lobato_peng.m in FP64, so they agree to FP64 rounding
(<= 1e-11 relative here); the numpy oracle itself is pinned in test_oracle.py.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import cpu_stem, multislice as oms, potential as opot
from totem_amd import optics, specimen, stem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module", autouse=True)
def built():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def _rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


@pytest.mark.parametrize("n", [8, 64, 320, 768])
def test_fft2_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    assert _rel(cpu_stem.fft2(x, -1), np.fft.fft2(x)) < 1e-13
    assert _rel(cpu_stem.fft2(x, +1), np.fft.ifft2(x) * n * n) < 1e-13


def _case(flag, ions=False, absorption=False, S=4, np_side=64):
    a = 3.905
    x = specimen.perovskite(2, 2, S // 2, a=a)
    rng = np.random.default_rng(5)
    x[:, 1:3] += rng.normal(0, 0.07, (len(x), 2))
    x[::5, 7] = 0.8
    if ions:
        o = np.nonzero(x[:, 0] == 8)[0][::3]
        x[o, 4] = -2
        x[o, 5] = 0.5
    dz = a / 2
    cfg = stem.StemConfig(atoms=x, np_side=np_side, dx=2 * a / np_side, n=32, kv=300.0, alpha_mrad=24.0,
                          slice_thickness=dz, n_slices=S, slice_origin=-dz / 4, paraflag=flag,
                          detectors_mrad=((20.0, 50.0), (0.0, 10.0)), width_red=3, hight_red=2, step=5,
                          begin_row=3, begin_col=4, mid_slices=(2,), focal_members=3, focal_rms_nm=3.0,
                          aberrations=optics.Aberrations(C1=-5.0))
    inp = stem.prepare(cfg)
    absorp = None
    if absorption:
        absorp = np.tile(np.array([0.01, 2.0, 0.02, 5.0, 0.005, 20.0, 0.0, 1.0, 0.0, 1.0]), (len(inp.ele_n), 1))
    return inp, absorp


def _args(inp, absorp):
    c = inp.cfg
    kw = {}
    if c.paraflag == "n":
        kw = dict(ele_n_corr=inp.ele_n_corr, corr_matrix=inp.corr_matrix, series_n_corr=inp.series_n_corr)
    ele_i = inp.ele_n_i if len(inp.ele_n_i) else None
    ser_i = inp.series_n_i if len(inp.ele_n_i) else None
    return (c.np_side, inp.ele_n, absorp, ele_i, None, inp.series_n, ser_i, inp.s2, inp.gx, inp.gy, inp.sigma,
            inp.parameter, inp.aper_true, c.paraflag), kw


@pytest.mark.parametrize("flag,ions,absorption", [("l", False, False), ("p", True, True), ("n", False, False),
                                                  ("l", True, True)])
def test_potential_matches_numpy_oracle(flag, ions, absorption):
    inp, absorp = _case(flag, ions, absorption)
    if ions:
        assert len(inp.ele_n_i) > 0
    args, kw = _args(inp, absorp)
    V0, t0 = opot.potential_slices(*args, **kw)
    V1, t1 = cpu_stem.potential_slices(*args, **kw, threads=2)
    assert _rel(V1, V0) < 1e-11
    assert _rel(t1, t0) < 1e-11
    # a slice range builds only those slices, identically
    _, tp = cpu_stem.potential_slices(*args, **kw, slices=(1, 3), threads=1)
    np.testing.assert_array_equal(tp[1:3], t1[1:3])
    assert not tp[0].any() and not tp[3].any()


@pytest.mark.parametrize("periodic", [False, True])
def test_stem_matches_numpy_oracle(periodic):
    inp, _ = _case("n")
    c = inp.cfg
    args, kw = _args(inp, None)
    _, t = opot.potential_slices(*args, **kw)
    ash = np.fft.ifftshift(inp.aper, axes=(-2, -1))
    r0, c0 = [], []
    for d in range(c.depth):
        r0.append(c.begin_row - 1 + (d // c.width_red) * c.step + (40 if periodic else 0))
        c0.append(c.begin_col - 1 + (d % c.width_red) * c.step + (45 if periodic else 0))
    out, mid = cpu_stem.stem_positions(t, inp.probe, inp.gfsf, inp.psf, ash, r0, c0, midslice=c.mid_slices,
                                       periodic=periodic, threads=3)
    assert len(inp.gfsf) == 3
    refs = [oms.stem_position(t, inp.probe, inp.gfsf, inp.psf, ash, r0[d], c0[d], midslice=c.mid_slices,
                              periodic=periodic) for d in range(c.depth)]
    ref = np.array([r for r, _ in refs])
    rmid = np.array([m for _, m in refs])
    # per detector, relative to its largest value (both detectors lie inside this grid's
    # 54 mrad band limit)
    assert ref[:, 0].min() > 1e-4
    for a in range(ref.shape[1]):
        scale = max(np.abs(ref[:, a]).max(), 1e-300)
        assert np.max(np.abs(out[:, a] - ref[:, a])) <= 1e-11 * scale
        scale = max(np.abs(rmid[:, :, a]).max(), 1e-300)
        assert np.max(np.abs(mid[:, :, a] - rmid[:, :, a])) <= 1e-11 * scale
    assert ref[:, 1].min() > 0.1                 # the BF disc carries most of the probe


def test_stem_refuses_windows_outside_the_potential():
    inp, _ = _case("l")
    t = np.ones((2, 64, 64), dtype=np.complex128)
    ash = np.fft.ifftshift(inp.aper, axes=(-2, -1))
    with pytest.raises(ValueError):
        cpu_stem.stem_positions(t, inp.probe, inp.gfsf, inp.psf, ash, [40], [0], threads=1)
