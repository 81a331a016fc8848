"""Pins of the FP64 oracle against closed forms, brute force and invariants (CPU only).

Each check is chosen so that a plausible restatement mistake fails it:
  * potential: Peng atom == sum of Gaussians; Lobato atom == K0/K1 closed form (first-order
    convergence in dx); DC identity; integer translation; Peng + (Lobato - Peng) == Lobato;
    atom position/phase sign and x<->column convention (off-grid atom position).
  * multislice: brute-force DFT matrices on 8x8; vacuum specimen == one-shot closed form;
    full-band bright field == 1; mid-slice k == exit face of a k-slice specimen; focal-series
    linearity; scan index <-> window translation.
"""
import numpy as np
import pytest
from scipy.special import k0, k1

from oracle import multislice as oms
from oracle import potential as opot
from totem_amd import optics, specimen, tables

C_POT = 47.87298  # 2 pi hbar^2/(e m) in V*Angstrom^2 with the reference's constants


def _single_atom(flag, Z, np_side, dx, x0, y0, B=0.0):
    gx, gy, s2 = optics.reciprocal_grids(np_side, dx)
    par = optics.potential_parameter(np_side, dx)
    if flag == "l":
        p = tables.lobato_table()[Z - 1]
        row = np.concatenate([[x0, y0, 1.0], tables.interleaved(p), [B]])
    else:
        p = tables.peng_table()[Z - 1]
        row = np.concatenate([[x0, y0, 1.0], tables.interleaved(p)])
    Vk = opot.reciprocal_potential(np_side, row[None, :], None, np.zeros((0, 26)), None, np.array([1.0]),
                                   np.array([]), s2, gx, gy, par, flag)
    V = opot.real_space_potential(Vk, np.ones_like(s2))[0]
    yy, xx = np.mgrid[0:np_side, 0:np_side] * dx
    return V, np.hypot(xx - x0, yy - y0), p


def test_parameter_constant():
    assert optics.potential_parameter(1, 1.0) == pytest.approx(C_POT, rel=2e-7)


def test_peng_atom_matches_gaussian_closed_form():
    # V(r) = C sum_i a_i (4 pi / b_i) exp(-4 pi^2 r^2 / b_i)   (f(s) = sum a exp(-b s^2), s = g/2)
    V, r, p = _single_atom("p", 8, 1024, 0.01, 5.133, 4.871)
    a, b = p[:5], p[5:]
    Va = C_POT * np.sum(a * 4 * np.pi / b * np.exp(-4 * np.pi ** 2 * r[..., None] ** 2 / b), axis=-1)
    m = r < 3.0
    err = np.abs(V.real - Va)[m].max() / Va[m].max()
    assert err < 2e-6, err
    # a swapped x/y or a flipped phase sign puts the atom elsewhere
    iy, ix = np.unravel_index(np.argmax(V.real), V.shape)
    assert abs(ix * 0.01 - 5.133) < 0.011 and abs(iy * 0.01 - 4.871) < 0.011


def test_lobato_atom_converges_to_bessel_closed_form():
    # V(r) = C sum a_i [ (2 pi/b_i) K0(2 pi r/sqrt b_i) + (2 pi^2 r / b_i^1.5) K1(2 pi r/sqrt b_i) ]
    errs = []
    for np_side, dx in ((512, 0.02), (1024, 0.01)):
        V, r, p = _single_atom("l", 38, np_side, dx, 5.133, 5.113)
        a, b = p[:5], p[5:]
        rr = np.maximum(r, 1e-9)[..., None]
        q = 2 * np.pi * rr / np.sqrt(b)
        Va = C_POT * np.sum(a * (2 * np.pi / b * k0(q) + 2 * np.pi ** 2 * rr / b ** 1.5 * k1(q)), axis=-1)
        m = (r >= 0.3) & (r < 3.0)
        errs.append(np.abs(V.real - Va)[m].max() / np.abs(Va[m]).max())
    assert errs[1] < 0.02, errs
    assert errs[0] / errs[1] > 1.7, errs  # the band-limit truncation error is first order in dx


def test_potential_dc_identity():
    # sum_r V(r) dx^2 = C * sum_j occ_j f_j(0)
    dx = 0.05
    V, _, p = _single_atom("l", 22, 128, dx, 1.7, 2.9)
    assert V.real.sum() * dx * dx == pytest.approx(C_POT * 2 * p[:5].sum(), rel=1e-6)
    V, _, p = _single_atom("p", 22, 128, dx, 1.7, 2.9)
    assert V.real.sum() * dx * dx == pytest.approx(C_POT * p[:5].sum(), rel=1e-6)


def test_potential_integer_translation():
    dx = 0.06
    V1, _, _ = _single_atom("l", 14, 64, dx, 1.23, 2.01)
    V2, _, _ = _single_atom("l", 14, 64, dx, 1.23 + 5 * dx, 2.01 - 3 * dx)
    np.testing.assert_allclose(V2, np.roll(np.roll(V1, 5, axis=1), -3, axis=0), atol=1e-9 * np.abs(V1).max())


def test_lobato_debye_waller_factor():
    # exp(-B s^2) on Lobato rows: B=0 vs B>0 differ exactly by that factor in reciprocal space
    np_side, dx = 64, 0.08
    gx, gy, s2 = optics.reciprocal_grids(np_side, dx)
    par = optics.potential_parameter(np_side, dx)
    p = tables.interleaved(tables.lobato_table()[37])
    rows = [np.concatenate([[1.0, 2.0, 0.7], p, [B]])[None, :] for B in (0.0, 0.6)]
    V0, V1 = [opot.reciprocal_potential(np_side, r, None, np.zeros((0, 26)), None, np.array([1.0]), np.array([]),
                                        s2, gx, gy, par, "l")[0] for r in rows]
    np.testing.assert_allclose(V1, V0 * np.exp(-0.6 * s2), rtol=1e-12, atol=1e-14)


# Absorptive coefficients (C_q, D_q interleaved as absorp_n rows hold them); synthetic values
# of the magnitude the reference's fitted tables have, D_q in Angstrom^2.
ABSORP = np.array([0.021, 0.45, 0.048, 1.9, 0.093, 7.5, 0.031, 21.0, 0.012, 63.0])


def _gauss_closed(c, r):
    """Inverse transform of sum_q c_2q exp(-c_2q+1 s^2): C sum_q c_2q (4 pi/c_2q+1) exp(-4 pi^2 r^2/c_2q+1)."""
    a, b = np.asarray(c)[0::2], np.asarray(c)[1::2]
    return C_POT * np.sum(a * 4 * np.pi / b * np.exp(-4 * np.pi ** 2 * r[..., None] ** 2 / b), axis=-1)


def _atom_potential(flag, np_side, dx, ele_n=None, absorp_n=None, ele_n_i=None, absorp_n_i=None):
    """Real-space potential of one slice holding the given rows (complex; centred k grid)."""
    gx, gy, s2 = optics.reciprocal_grids(np_side, dx)
    par = optics.potential_parameter(np_side, dx)
    lob = flag == "l"
    e = np.zeros((0, 14 if lob else 13)) if ele_n is None else np.atleast_2d(ele_n)
    ei = np.zeros((0, 26 if lob else 25)) if ele_n_i is None else np.atleast_2d(ele_n_i)
    ab = None if absorp_n is None else np.atleast_2d(absorp_n)
    abi = None if absorp_n_i is None else np.atleast_2d(absorp_n_i)
    Vk = opot.reciprocal_potential(np_side, e, ab, ei, abi, np.array([len(e)], float), np.array([len(ei)], float),
                                   s2, gx, gy, par, flag)
    return opot.real_space_potential(Vk, np.ones_like(s2))[0]


def _radius(np_side, dx, x0, y0):
    yy, xx = np.mgrid[0:np_side, 0:np_side] * dx
    return np.hypot(xx - x0, yy - y0)


def _rel_max(V, Va, m):
    return np.abs(V - Va)[m].max() / np.abs(Va[m]).max()


def test_peng_absorptive_part_matches_gaussian_closed_form():
    # ..._peng.m:41-53: V_k = P occ (sum a e^{-b s^2} + i sum C e^{-D s^2}) phase, so Re V is the
    # elastic sum of Gaussians and Im V the absorptive one
    np_side, dx, x0, y0 = 1024, 0.01, 5.133, 4.871
    p = tables.interleaved(tables.peng_table()[37])
    V = _atom_potential("p", np_side, dx, ele_n=np.concatenate([[x0, y0, 0.8], p]), absorp_n=ABSORP)
    r = _radius(np_side, dx, x0, y0)
    m = r < 3.0
    assert _rel_max(V.real, 0.8 * _gauss_closed(p, r), m) < 2e-6
    assert _rel_max(V.imag, 0.8 * _gauss_closed(ABSORP, r), m) < 2e-6


def test_lobato_debye_waller_leaves_absorptive_part_undamped():
    # ..._peng.m:120-133: exp(-B s^2) multiplies the Lobato elastic sum only; the absorptive
    # Gaussians enter undamped, and they add nothing to Re V
    np_side, dx, x0, y0, B = 1024, 0.01, 4.92, 5.31, 0.5
    p = tables.interleaved(tables.lobato_table()[37])
    row = np.concatenate([[x0, y0, 1.0], p, [B]])
    V = _atom_potential("l", np_side, dx, ele_n=row, absorp_n=ABSORP)
    V0 = _atom_potential("l", np_side, dx, ele_n=row)
    r = _radius(np_side, dx, x0, y0)
    m = r < 3.0
    assert _rel_max(V.imag, _gauss_closed(ABSORP, r), m) < 2e-6
    np.testing.assert_allclose(V.real, V0.real, rtol=0, atol=1e-12 * np.abs(V0.real).max())


def test_fully_ionic_peng_atom_matches_gaussian_closed_form():
    # ionicity 1 takes the Peng ion row of (Z, valence) (getpotentialpara_lotabo_peng.m:35-53)
    np_side, dx, x0, y0 = 1024, 0.01, 5.07, 4.96
    x = np.array([[38, x0, y0, 1.0, 2, 1.0, 0.0, 1.0]])
    all_n, all_n_i = specimen.atom_parameters(x, "p")
    ele_n, _, sn, _ = specimen.slice_atoms(all_n, all_n_i, 2.0, np.array([0.0]), 0, "p", 0.0)
    assert sn.tolist() == [1.0]
    V = _atom_potential("p", np_side, dx, ele_n=ele_n)
    ion = tables.peng_ion_table()
    c_ion = tables.interleaved(ion[(ion[:, 0] == 38) & (ion[:, 1] == 2)][0, 2:])
    r = _radius(np_side, dx, x0, y0)
    m = r < 3.0
    assert _rel_max(V.real, _gauss_closed(c_ion, r), m) < 2e-6
    # the neutral atom's potential is measurably different, so the check above discriminates
    assert _rel_max(_gauss_closed(tables.interleaved(tables.peng_table()[37]), r), _gauss_closed(c_ion, r), m) > 1e-2


@pytest.mark.parametrize("w", [0.3, 0.75])
def test_mixed_ionicity_is_linear_in_the_ion_weight(w):
    # ..._peng.m:58-99: occ (w f_ion + (1 - w) f_atom); the ion and atom slots of an ele_n_i row
    # (getpotentialpara_lotabo_peng.m:59-86) must carry the ion and neutral coefficients
    np_side, dx, x0, y0 = 1024, 0.01, 5.11, 5.02
    Vs = {}
    for ionicity in (w, 1.0, 0.0):
        x = np.array([[22, x0, y0, 1.0, 4, ionicity, 0.0, 0.9]])
        all_n, all_n_i = specimen.atom_parameters(x, "p")
        ele_n, ele_n_i, sn, sni = specimen.slice_atoms(all_n, all_n_i, 2.0, np.array([0.0]), 0, "p", 0.0)
        Vs[ionicity] = _atom_potential("p", np_side, dx, ele_n=ele_n if len(ele_n) else None,
                                       ele_n_i=ele_n_i if len(ele_n_i) else None)
    # ionicity 1 and 0 give single-form rows of occupancy 0.9 (ion and neutral coefficients),
    # so V(w) = w V(1) + (1 - w) V(0); swapped weights or slots break it
    ref = w * Vs[1.0] + (1 - w) * Vs[0.0]
    np.testing.assert_allclose(Vs[w], ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    ion = tables.peng_ion_table()
    c_ion = tables.interleaved(ion[(ion[:, 0] == 22) & (ion[:, 1] == 4)][0, 2:])
    c_atom = tables.interleaved(tables.peng_table()[21])
    r = _radius(np_side, dx, x0, y0)
    m = r < 3.0
    closed = 0.9 * (w * _gauss_closed(c_ion, r) + (1 - w) * _gauss_closed(c_atom, r))
    assert _rel_max(Vs[w].real, closed, m) < 2e-6


def test_lobato_ion_rows_weight_both_sets_and_damp_them_by_b():
    # ..._peng.m:141-177: occ ((w1 L1 + w2 L2) e^{-B s^2} + i sum C e^{-D s^2})
    np_side, dx, x0, y0, B = 256, 0.04, 4.1, 5.3, 0.4
    L1 = tables.interleaved(tables.lobato_table()[21])
    L2 = tables.interleaved(tables.lobato_table()[7])
    single = [_atom_potential("l", np_side, dx, ele_n=np.concatenate([[x0, y0, 0.7], L, [B]])) for L in (L1, L2)]
    for w1, w2 in ((1.0, 0.0), (0.0, 1.0), (0.35, 0.65)):
        row = np.concatenate([[x0, y0, 0.7, w1], L1, [w2], L2, [B]])
        V = _atom_potential("l", np_side, dx, ele_n_i=row)
        ref = w1 * single[0] + w2 * single[1]
        np.testing.assert_allclose(V, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    # absorption on an ion row: Im V is the undamped Gaussian sum (fine grid: no Nyquist leak)
    np_side, dx = 1024, 0.01
    row = np.concatenate([[x0, y0, 0.7, 0.35], L1, [0.65], L2, [B]])
    V = _atom_potential("l", np_side, dx, ele_n_i=row, absorp_n_i=ABSORP)
    r = _radius(np_side, dx, x0, y0)
    assert _rel_max(V.imag, 0.7 * _gauss_closed(ABSORP, r), r < 3.0) < 2e-6


def test_peng_plus_correction_equals_lobato():
    a = 3.905
    x = specimen.perovskite(2, 2, 1, a=a)
    np_side, dx = 64, 2 * a / 64
    st = np.array([-0.5, a / 2 - 0.5])
    gx, gy, s2 = optics.reciprocal_grids(np_side, dx)
    par = optics.potential_parameter(np_side, dx)
    res = {}
    for flag in ("l", "n"):
        all_n, all_n_i = specimen.atom_parameters(x, flag)
        ele_n, ele_n_i, sn, sni = specimen.slice_atoms(all_n, all_n_i, a / 2, st, 0, flag, 0.0)
        kw = {}
        if flag == "n":
            cn = specimen.corr_parameters(x)
            ec, ci, sc = specimen.slice_corr(cn, a / 2, st, 0)
            kw = dict(ele_n_corr=ec, corr_matrix=specimen.corr_info_matrix(ci, s2), series_n_corr=sc)
        res[flag] = opot.reciprocal_potential(np_side, ele_n, None, ele_n_i, None, sn, sni, s2, gx, gy, par, flag,
                                              **kw)
    np.testing.assert_allclose(res["n"], res["l"], rtol=1e-10, atol=1e-10 * np.abs(res["l"]).max())


def test_transmission_of_vacuum_and_real_potential():
    V = np.zeros((2, 8, 8), dtype=complex)
    np.testing.assert_array_equal(opot.transmission(V, 0.73), np.ones((2, 8, 8)))
    V = np.random.default_rng(1).standard_normal((8, 8)) * 100
    t = opot.transmission(V.astype(complex), optics.interaction_sigma(200.0))
    np.testing.assert_allclose(np.abs(t), 1.0, atol=1e-14)
    np.testing.assert_allclose(np.angle(t), np.angle(np.exp(1j * V * 0.7288688742821855e-3)), atol=1e-12)


# ---------------------------------------------------------------------------- multislice

def _rand_complex(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def test_multislice_matches_bruteforce_dft():
    rng = np.random.default_rng(7)
    n, np_side, S = 8, 12, 3
    t = np.exp(1j * rng.standard_normal((S, np_side, np_side)) - 0.1 * rng.random((S, np_side, np_side)))
    probe = _rand_complex(rng, (1, n, n))
    psf = np.exp(1j * rng.standard_normal((n, n)))
    psf[0, 3] = 0
    aper = (rng.random((2, n, n)) > 0.5).astype(float)
    r0, c0 = 2, 3
    res, _ = oms.stem_position(t, probe, np.ones(1), psf, np.fft.ifftshift(aper, axes=(-2, -1)), r0, c0)
    W = np.exp(-2j * np.pi * np.outer(np.arange(n), np.arange(n)) / n)
    Wi = np.conj(W) / n
    w = probe[0].copy()
    for k in range(S):
        w = w * t[k, r0:r0 + n, c0:c0 + n]
        w = W @ w @ W.T
        w = w * psf
        w = Wi @ w @ Wi.T
    w = W @ w @ W.T
    ws = np.roll(np.roll(w, n // 2, 0), n // 2, 1)  # fftshift, compared with centred masks
    ref = [np.sum(np.abs(ws) ** 2 * aper[a]) for a in range(2)]
    np.testing.assert_allclose(res, ref, rtol=1e-12)


def _vacuum_case(n=32, S=5):
    dx, kv = 0.2, 200.0
    probe, gfsf = optics.probe_stack(n, dx, kv, 15.0)
    psf = optics.propagator(n, dx, kv, 2.0)
    aper = optics.detector_masks(n, dx, kv, [(0.0, 12.0), (10.0, 60.0), (0.0, 400.0)])
    t = np.ones((S, n, n), dtype=complex)
    return probe, gfsf, psf, aper, t


def test_vacuum_closed_form_and_bright_field_normalisation():
    probe, gfsf, psf, aper, t = _vacuum_case()
    n = probe.shape[-1]
    res, _ = oms.stem_cpu(t, probe, gfsf, psf, 1, 1, 1, 1, 1, aper)
    P = np.fft.fftshift(np.fft.fft2(probe[0]))
    band = np.fft.fftshift(np.abs(psf) > 0)
    closed = [np.sum(np.abs(P) ** 2 * band * aper[a]) for a in range(3)]
    np.testing.assert_allclose(res[0], closed, rtol=1e-12)
    # a detector covering the whole probe aperture (inside the band limit) collects all of it
    assert res[0, 2] / n ** 2 == pytest.approx(1.0, rel=1e-12)


def test_mid_slice_equals_exit_of_truncated_specimen():
    rng = np.random.default_rng(3)
    n, S, k = 16, 6, 4
    t = np.exp(1j * 0.5 * rng.standard_normal((S, 24, 24)))
    probe, gfsf = optics.probe_stack(n, 0.1, 200.0, 20.0)
    psf = optics.propagator(n, 0.1, 200.0, 1.5)
    aper = optics.detector_masks(n, 0.1, 200.0, [(0, 20), (30, 200)])
    res, mid = oms.stem_cpu(t, probe, gfsf, psf, 3, 5, 2, 2, 4, aper, midslice=(k,))
    res_k, _ = oms.stem_cpu(t[:k], probe, gfsf, psf, 3, 5, 2, 2, 4, aper)
    D, A = res.shape
    np.testing.assert_allclose(mid[:, 0].reshape(A, D).T, res_k, rtol=1e-11)


def test_focal_series_is_linear_in_gfsf():
    rng = np.random.default_rng(5)
    n = 16
    t = np.exp(1j * 0.3 * rng.standard_normal((3, 20, 20)))
    probes = _rand_complex(rng, (2, n, n))
    psf = optics.propagator(n, 0.1, 200.0, 1.0)
    aper = optics.detector_masks(n, 0.1, 200.0, [(0, 300)])
    g = np.array([0.3, 0.7])
    r2, _ = oms.stem_cpu(t, probes, g, psf, 1, 1, 2, 2, 2, aper)
    ra, _ = oms.stem_cpu(t, probes[:1], np.ones(1), psf, 1, 1, 2, 2, 2, aper)
    rb, _ = oms.stem_cpu(t, probes[1:], np.ones(1), psf, 1, 1, 2, 2, 2, aper)
    np.testing.assert_allclose(r2, 0.3 * ra + 0.7 * rb, rtol=1e-12)


def test_scan_index_is_window_translation():
    rng = np.random.default_rng(9)
    n, step = 16, 3
    t = np.exp(1j * 0.4 * rng.standard_normal((2, 30, 30)))
    probe, g = optics.probe_stack(n, 0.1, 200.0, 20.0)
    psf = optics.propagator(n, 0.1, 200.0, 1.0)
    aper = optics.detector_masks(n, 0.1, 200.0, [(30, 200)])
    res, _ = oms.stem_cpu(t, probe, g, psf, 1, 1, 3, 2, step, aper)
    # d = i + j*width_red: column index i moves the window right, row index j moves it down
    shifted_c = np.roll(t, -step, axis=2)
    shifted_r = np.roll(t, -step, axis=1)
    r_c, _ = oms.stem_cpu(shifted_c, probe, g, psf, 1, 1, 3, 2, step, aper)
    r_r, _ = oms.stem_cpu(shifted_r, probe, g, psf, 1, 1, 3, 2, step, aper)
    np.testing.assert_allclose(res[1], r_c[0], rtol=1e-12)
    np.testing.assert_allclose(res[3], r_r[0], rtol=1e-12)
