"""CPU tests of the host side: tables, optics set-up, specimen preparation, the C-ABI
library's exports and the mex gateway's argument mapping.  No GPU calls."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from totem_amd import optics, specimen, stem, tables

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# --------------------------------------------------------------------------- tables

def test_table_shapes_and_known_values():
    assert tables.lobato_table().shape == (103, 10)
    assert tables.peng_table().shape == (98, 10)
    # f_Sr(0): Lobato 2*sum(a) and Peng sum(a), both ~13.11 Angstrom
    assert 2 * tables.lobato_table()[37, :5].sum() == pytest.approx(13.10990, abs=1e-5)
    assert tables.peng_table()[37, :5].sum() == pytest.approx(13.1088, abs=1e-4)


@pytest.mark.parametrize("Z", [6, 8, 14, 22, 38, 79])
def test_lobato_and_peng_tables_agree(Z):
    # two independent parameterisations of the same scattering factor (transcription check)
    s = np.linspace(0, 2.0, 41)
    fl = tables.lobato_f(Z, 4 * s * s)
    fp = tables.peng_f(Z, s * s)
    assert np.max(np.abs(fl - fp) / fl.max()) < 0.05


# --------------------------------------------------------------------------- optics

def test_wavelength_and_sigma_closed_forms():
    # Kirkland, Advanced Computing in EM (sigma vs voltage), as ToTEM_submit.m:1490-1493 cross-checks
    assert optics.wavelength_A(200) == pytest.approx(0.025078, abs=2e-6)
    assert optics.wavelength_A(300) == pytest.approx(0.019687, abs=2e-6)
    assert optics.interaction_sigma(200) == pytest.approx(0.728869, abs=2e-6)
    assert optics.interaction_sigma(300) == pytest.approx(0.652640, abs=2e-6)


def test_propagator_band_limit_and_unit_modulus():
    n, dx = 64, 0.1
    h = optics.propagator(n, dx, 200.0, 2.0)
    f = np.fft.ifftshift(optics.centred_freq(n, dx))
    kx, ky = np.meshgrid(f, f)
    inside = kx ** 2 + ky ** 2 <= 1 / (9 * dx * dx)
    np.testing.assert_allclose(np.abs(h[inside]), 1.0, atol=1e-14)
    assert np.all(h[~inside] == 0)
    # Fresnel composition: H(dz1) H(dz2) = H(dz1 + dz2) inside the band
    h2 = optics.propagator(n, dx, 200.0, 3.0)
    h5 = optics.propagator(n, dx, 200.0, 5.0)
    np.testing.assert_allclose((h * h2)[inside], h5[inside], atol=1e-12)


def test_detector_masks_and_probe_normalisation():
    n, dx = 128, 0.08
    aper = optics.detector_masks(n, dx, 200.0, [(60, 200), (0, 10)])
    lam = optics.wavelength_A(200)
    f = optics.centred_freq(n, dx)
    kx, ky = np.meshgrid(f, f)
    k = np.sqrt(kx ** 2 + ky ** 2) * lam * 1e3
    assert aper[0][n // 2, n // 2] == 0 and aper[1][n // 2, n // 2] == 1
    assert np.all(aper[0][(k > 61) & (k < 199)] == 1)
    p, g = optics.probe_stack(n, dx, 200.0, 20.0, optics.Aberrations(C1=-5, Cs_um=1.0), mtotal=3, focal_rms_nm=4)
    np.testing.assert_allclose(np.sum(np.abs(p) ** 2, axis=(1, 2)), 1.0, rtol=1e-12)
    assert len(g) == 3 and g[0] == pytest.approx(g[2]) and g[1] > g[0]
    # the probe's spectrum is confined to the aperture
    P = np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(p[1])))
    out = k > 20.0 + 1e-9
    assert np.max(np.abs(P[out])) < 1e-12 * np.max(np.abs(P))


# ------------------------------------------------------------------------- specimen

def test_perovskite_counts_and_slicing():
    a = 3.905
    x = specimen.perovskite(5, 5, 10, a=a)
    assert x.shape == (1250, 8)
    assert sorted(np.unique(x[:, 0])) == [8, 22, 38]
    cfg = stem.perovskite_c1()
    inp = stem.prepare(cfg)
    assert inp.ele_n.shape == (1250, 14)
    np.testing.assert_array_equal(inp.series_n, [50, 75] * 10)
    # columns: x y occ a1 b1 ... a5 b5 B (Lobato); coefficients are the table's
    sr = inp.ele_n[inp.ele_n[:, 3] == tables.lobato_table()[37, 0]]
    assert len(sr) == 250
    np.testing.assert_array_equal(sr[0, 3:13], tables.interleaved(tables.lobato_table()[37]))


def test_multiceng_weights_sum_to_one():
    # an atom split over neighbouring slices keeps its total weight (erf contributions)
    a = 3.905
    x = specimen.perovskite(1, 1, 3, a=a)
    x[:, 6] = 0.5
    all_n, all_n_i = specimen.atom_parameters(x, "p")
    st = np.arange(6) * a / 2 - 0.3
    ele_n, _, sn, _ = specimen.slice_atoms(all_n, all_n_i, a / 2, st, 1, "p", 0.0)
    assert len(ele_n) > len(all_n)
    # the weighted a1 of every atom column (same x, y) summed over slices equals its atoms' a1 sum
    got, ref = {}, {}
    for r in ele_n:
        k = (round(r[0], 9), round(r[1], 9))
        got[k] = got.get(k, 0.0) + r[3]
    for r in all_n:
        k = (round(r[1], 9), round(r[2], 9))
        ref[k] = ref.get(k, 0.0) + r[6]
    assert got.keys() == ref.keys()
    for k in ref:
        assert got[k] == pytest.approx(ref[k], rel=1e-9)


def test_corr_slicing_classes():
    x = specimen.perovskite(2, 2, 2)
    cn = specimen.corr_parameters(x)
    st = np.arange(4) * 3.905 / 2 - 0.4
    ec, ci, sc = specimen.slice_corr(cn, 3.905 / 2, st, 0)
    assert ci.shape == (3, 20) and ec.shape[1] == 5
    assert set(np.unique(ec[:, 4])) == {1.0, 2.0, 3.0}
    assert sc.sum() == len(x)
    # sorted unique rows: MATLAB unique(...,'rows') orders lexicographically
    assert np.all(np.diff(ci[:, 0]) > 0)


def test_frozen_phonon_statistics_and_slab():
    rng = np.random.default_rng(11)
    x = specimen.perovskite(10, 10, 4)
    all_n, _ = specimen.atom_parameters(x, "p")
    B = 8 * np.pi ** 2 * 0.08 ** 2
    all_n[:, 4] = B
    st = np.arange(8) * 3.905 / 2 - 0.5
    out = specimen.frozen_phonon(all_n, 3.905 / 2, st, (), rng)
    d = out[:, 1:3] - all_n[:, 1:3]
    assert np.std(d) == pytest.approx(0.08, rel=0.05)
    assert np.all(out[:, 3] >= st[0]) and np.all(out[:, 3] < st[-1] + 3.905 / 2)


def test_mex_args_layout():
    cfg = stem.perovskite_c1(n=64, np_side=256, scan=2, step=3)
    cfg.mid_slices = (5,)
    inp = stem.prepare(cfg)
    args = stem.mex_args(inp)
    assert len(args) == 41
    assert args[0].shape == (1, 64, 64) and args[2].shape == (1, 64, 64)
    assert args[9].shape == (256, 256) and np.all(args[9][0] == args[9][5])  # kx varies along columns
    assert args[32].dtype == np.int32 and list(args[32]) == [5]


# -------------------------------------------------------------------------- C ABI

def _header_functions():
    src = open(os.path.join(ROOT, "include", "totem_abi.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = set(re.findall(r"^\s*(?:const\s+)?(?:void|int|size_t|char)\s*\*?\s*(\w+)\s*\(", src, flags=re.M))
    return names


def _exports(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    return {l.split()[-1] for l in out.splitlines() if " T " in l}


def test_library_exports_every_declared_symbol():
    from totem_amd import engine
    names = _header_functions()
    assert {"myProjectfun", "totem_project_lobato", "totem_stem_run", "totem_build_transmission"} <= names
    core = _exports(engine.CORE)
    assert names - {"myProjectfun"} <= core, names - core
    for lib in engine.VARIANT_LIBS.values():
        assert "myProjectfun" in _exports(os.path.join(engine.LIB_DIR, lib))
    assert set(engine.EXPORTS) == names


def test_library_loads_without_gpu():
    from totem_amd import engine
    L = ctypes.CDLL(engine.CORE)
    L.totem_abi_version.restype = ctypes.c_int
    assert L.totem_abi_version() == engine.ABI_VERSION
    for n in (64, 128, 256, 320, 512, 768, 1024, 1536, 2048):
        assert L.totem_fft_size_supported(n) == 1
    assert L.totem_fft_size_supported(1000) == 0


def test_mex_gateway_argument_mapping(tmp_path):
    exe = tmp_path / "drive"
    subprocess.run(["g++", "-O0", "-I" + os.path.join(ROOT, "tests", "mexstub"), "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "mexstub", "drive.cpp"),
                    os.path.join(ROOT, "totem_amd", "mex", "waveCuda.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    assert "H=64 W=64 D=12 PH=256 PW=256 S=9 step=3 br=10 bc=20 wr=4 wn=3 an=2 cl=1 cim=0" in out
    assert "absorp_null=1 absorp_i_null=0 corr_null=1 w0=0.25 aper2=0.5 mid=5" in out
    assert "out0=12x2:42 out1=24x1 out2=65536x9:7" in out
    assert "error=\n" in out
    assert "nrhs40=Invalid number of input arguments" in out
    out = subprocess.run([str(exe), "fail"], capture_output=True, text=True, check=True).stdout
    assert "error=engine failure" in out


def test_diamond_110_repeat_and_bond_length():
    from scipy.spatial import cKDTree
    from totem_amd import specimen
    a = 5.431
    x = specimen.diamond_110(10 * a / np.sqrt(2), 7 * a, 10 * a / np.sqrt(2), a=a)
    assert len(x) == 4 * 10 * 7 * 10                        # 4 sites per a/sqrt2 x a x a/sqrt2 repeat
    p = x[:, 1:4]
    d, _ = cKDTree(p).query(p, k=5)
    inner = np.all((p > 3) & (p < p.max(axis=0) - 3), axis=1)
    np.testing.assert_allclose(d[inner, 1:], a * np.sqrt(3) / 4, rtol=1e-9)   # 4 bonds of a*sqrt(3)/4


def test_amorphous_density_separation_and_seed():
    from scipy.spatial import cKDTree
    from totem_amd import specimen
    x = specimen.amorphous(30.0, 30.0, 5.0, 25.0, seed=7)
    assert len(x) == round(0.1003 * 30 * 30 * 20)
    p = x[:, 1:4]
    assert p[:, 2].min() >= 5.0 and p[:, 2].max() < 25.0 and p[:, :2].min() >= 0 and p[:, :2].max() < 30.0
    d, _ = cKDTree(p, boxsize=[30.0, 30.0, 1e3]).query(p, k=2)
    assert d[:, 1].min() >= 1.4
    np.testing.assert_array_equal(x, specimen.amorphous(30.0, 30.0, 5.0, 25.0, seed=7))
    assert not np.array_equal(x, specimen.amorphous(30.0, 30.0, 5.0, 25.0, seed=8))


def test_bench_reads_committed_traffic_profile():
    """bench.py takes the roofline's HBM traffic from the committed rocprof summary: both passes
    parse, with read + write within 10% of the algorithmic bytes of a 26-probe launch (the band
    columns of the headline propagator: 688 of 1024 with 4-column groups)."""
    import bench
    from totem_amd import engine, optics, stem
    cfg = stem.perovskite_c2(scan=16)
    psf = optics.propagator(cfg.n, cfg.dx, cfg.kv, cfg.slice_thickness)
    _, band = engine.stem_band(optics.unit_complex64(psf), 4)
    assert band == 688
    n2 = 1024.0 * 1024
    algs = {"k_pass_rows<1024>": 26 * 16.0 * band * 1024 + 8.0 * n2,
            "k_pass_cols<1024>": 26 * 16.0 * band * 1024 + 8.0 * band * 1024}
    for k, alg in algs.items():
        tr = bench.measured_traffic(k, 1024, "c2")
        assert tr is not None, k
        assert 0.95 * alg < tr < 1.15 * alg, (k, tr / alg)
    assert bench.measured_traffic("k_pass_rows<512>", 512, "c2") is None
    # and the VALU issue fraction of each pass from the committed SQ counter summary
    for k in algs:
        v = bench.measured_valu_issue(k, 1024, "c2")
        assert v is not None and 0.3 < v < 1.0, (k, v)


# --------------------------------------------------------------------------- potential planner

def _planner_case(multiceng=0, per_atom_dw=False, flag="p"):
    from dataclasses import replace
    a = 3.905
    x = specimen.perovskite(3, 3, 3, a=a)
    rng = np.random.default_rng(8)
    x[:, 3] += rng.uniform(-0.3, 0.3, len(x))     # z off the lattice planes: many atoms near faces
    dz = a / 2
    cfg = stem.StemConfig(atoms=x, np_side=128, dx=3 * a / 128, n=64, kv=300.0, alpha_mrad=24.0,
                          slice_thickness=dz, n_slices=6, slice_origin=-dz / 4, paraflag=flag, multiceng=multiceng)
    if per_atom_dw:
        x = x.copy()
        x[:, 6] = rng.uniform(0.2, 0.8, len(x))
        cfg = replace(cfg, atoms=x, dw_selected=True)
    return stem.prepare(cfg)


def _plan(inp):
    from totem_amd import engine
    c = inp.cfg
    return engine.potential_plan_stats(c.paraflag, c.np_side, c.n_slices, inp.series_n, inp.ele_n, inp.gx, inp.gy,
                                       inp.s2, inp.aper_true, inp.sigma, inp.parameter)


def test_planner_splits_erf_weighted_rows_into_fewer_classes():
    # multiceng = 1: rows near faces carry erf-scaled a_q (one coefficient set per row) but share
    # their element's b_q, so one class per (element, Gaussian term) replaces one per row
    inp = _planner_case(multiceng=1)
    rows = len(inp.ele_n)
    distinct_rows = len({tuple(r[3:]) for r in inp.ele_n})
    st = _plan(inp)
    assert distinct_rows > 16
    assert st["split"]
    assert st["classes"] <= 5 * 3 < distinct_rows        # Sr, Ti, O: five b_q each
    assert rows < st["terms"] <= 5 * rows
    # a static specimen sliced without erf weighting keeps one class per element
    st0 = _plan(_planner_case(multiceng=0))
    assert not st0["split"] and st0["classes"] == 3


def test_planner_reports_paired_frequencies():
    # the prepared grids come from integer multiples of 1/(np dx): k[np - i] = -k[i] exactly
    inp = _planner_case(multiceng=0)
    c = inp.cfg
    n = c.np_side
    assert np.array_equal(inp.gx[0, n - 1:0:-1], -inp.gx[0, 1:])
    assert _plan(inp)["paired"]
    # a grid offset by a fraction of a pixel has no pairs: the accumulation covers every tile
    from totem_amd import engine
    gx = inp.gx + 0.25 * (inp.gx[0, 1] - inp.gx[0, 0])
    st = engine.potential_plan_stats(c.paraflag, n, c.n_slices, inp.series_n, inp.ele_n, gx, inp.gy, inp.s2,
                                     inp.aper_true, inp.sigma, inp.parameter)
    assert not st["paired"]


def test_planner_keeps_rows_whose_gaussians_do_not_repeat():
    # per-atom Debye-Waller B folded into every b_q: splitting would make five classes per row
    inp = _planner_case(per_atom_dw=True)
    distinct_rows = len({tuple(r[3:]) for r in inp.ele_n})
    assert distinct_rows == len(inp.ele_n)
    st = _plan(inp)
    assert not st["split"]
    assert st["classes"] == distinct_rows and st["terms"] == len(inp.ele_n)


def _band_brute(flags, group):
    """Shortest length of a wrapped run of whole groups covering every flagged column."""
    n = len(flags)
    G = n // group
    on = [any(flags[g * group:(g + 1) * group]) for g in range(G)]
    if not any(on):
        return 0
    best = G
    for s in range(G):
        for ln in range(1, G + 1):
            if all(not on[g] or (g - s) % G < ln for g in range(G)):
                best = min(best, ln)
                break
    return best * group


def test_stem_band_is_the_shortest_wrapped_run_of_groups():
    """totem_stem_band (the columns a run processes): aligned to the group width, covers every
    column of prop holding a nonzero entry, and no shorter wrapped run of groups does; a zero
    propagator gives an empty band and one without an empty group the whole grid."""
    from totem_amd import engine
    rng = np.random.default_rng(11)
    n, group = 64, 4
    cases = [np.zeros(n, int), np.ones(n, int)]
    for _ in range(60):
        f = np.zeros(n, int)
        k = rng.integers(1, 5)
        for _ in range(k):  # a few wrapped runs of flagged columns
            a, ln = rng.integers(0, n), rng.integers(1, 24)
            f[(a + np.arange(ln)) % n] = 1
        cases.append(f)
    for f in cases:
        prop = np.zeros((3, n), np.complex64)
        prop[1, f.astype(bool)] = 0.5 - 0.25j
        lo, ln = engine.stem_band(prop, group)
        assert lo % group == 0 and ln % group == 0 and 0 <= lo < n and 0 <= ln <= n
        for c in np.flatnonzero(f):
            assert (c - lo) % n < ln, (f, lo, ln, c)
        assert ln == _band_brute(list(f), group), (f, lo, ln)
    # the propagator of a 2/3 band-limited grid: DC at column 0, the band wraps through it
    from totem_amd import optics
    psf = optics.propagator(96, 0.1, 200, 1.9525)
    lo, ln = engine.stem_band(psf, 4)
    cols = np.flatnonzero(np.abs(psf).max(axis=0) > 0)
    assert 0 < ln < 96 and lo > 0 and all((c - lo) % 96 < ln for c in cols)
