"""FP64 restatement of the reference CPU potential builder (oracle; test infrastructure only).

Follows ``GetPotential4AllSlice_multicore_lotabo_peng_corr.m`` (which contains
``GetPotential4AllSlice_multicore_lotabo_peng.m`` as its non-corrected branch):

* correction term first, ``..._corr.m:15-65``:
  V_k += PARAMETER * occ*w * C_cls(s) * exp(-B s^2) * exp(-2 pi i (gx x + gy y))
* Peng, ``GetPotential4AllSlice_multicore_lotabo_peng.m:25-99``:
  V_k += PARAMETER * occ * (sum a_i exp(-b_i s^2) [+ i * absorptive]) * phase
* Lobato, ``GetPotential4AllSlice_multicore_lotabo_peng.m:102-181``:
  V_k += PARAMETER * occ * (sum a_i (2+b_i g^2)/(1+b_i g^2)^2 * exp(-B s^2) [+ i * absorptive]) * phase,
  g^2 = 4 s^2
* then, per slice, ``..._peng.m:187-190``: V = ifft2(ifftshift(V_k * APERTURE)) * Np^2 and
  t = exp(i * V * sigma / 1000).

One deliberate difference: the MATLAB ion branch indexes the ion rows of slice i
with the neutral-atom offsets ``beg1`` (``..._peng.m:61,141``); the reference GPU
engine uses the ion running offset (``cuda_source_lobato.cu`` ``cal_p``,
``sumSeries_n_i``).  This oracle uses the ion offsets, i.e. the GPU engine's
semantics; the two agree whenever there are no mixed-ionicity atoms.

Pins (``tests/test_oracle.py``): DC identity sum_r V dx^2 = 47.87298 * sum occ f(0);
single Peng atom vs the closed-form projected potential (a sum of Gaussians);
single Lobato atom vs its closed form (K0/K1 Bessel terms) away from the core;
integer-pixel translation; Peng + (Lobato - Peng) correction == Lobato;
absorptive part: Im V of a Peng or Lobato atom (with and without ion rows) equals
the Gaussian closed form of its C/D coefficients, undamped by Lobato's B, and
leaves Re V unchanged; a fully ionic Peng atom equals the Gaussian closed form of
its (Z, valence) row; mixed ionicity is linear in the ion weight
(V(w) = w V(ion) + (1 - w) V(atom)) and matches the closed form; Lobato ion rows
weight both coefficient sets and damp both by B.
"""
import numpy as np


def _phase(gx, gy, x, y):
    return np.exp(-2j * np.pi * (gx * x + gy * y))


def _peng_sum(c, s2):
    return sum(c[2 * q] * np.exp(-s2 * c[2 * q + 1]) for q in range(5))


def _lobato_sum(c, g2):
    return sum(c[2 * q] * (2 + c[2 * q + 1] * g2) / (1 + c[2 * q + 1] * g2) ** 2 for q in range(5))


def _offsets(series):
    series = np.asarray(series, dtype=np.int64) if series is not None else np.zeros(0, np.int64)
    return np.concatenate([[0], np.cumsum(series)]).astype(np.int64)


def reciprocal_potential(np_side, ele_n, absorp_n, ele_n_i, absorp_n_i, series_n, series_n_i,
                         s2, gx, gy, parameter, paraflag,
                         ele_n_corr=None, ele_n_i_corr=None, corr_matrix=None,
                         series_n_corr=None, series_n_i_corr=None):
    """Per-slice reciprocal-space potential (centred k grid), complex128 [S, Np, Np]."""
    S = max(len(series_n) if series_n is not None else 0, len(series_n_i) if series_n_i is not None else 0)
    V = np.zeros((S, np_side, np_side), dtype=np.complex128)
    g2 = 4 * s2
    if paraflag == "n" and corr_matrix is not None and len(corr_matrix):
        b1 = _offsets(series_n_corr)
        for i in range(len(series_n_corr) if series_n_corr is not None else 0):
            for a in range(b1[i], b1[i + 1]):
                r = ele_n_corr[a]
                V[i] += parameter * r[3] * corr_matrix[int(r[4]) - 1] * np.exp(-s2 * r[2]) * _phase(gx, gy, r[0], r[1])
        b2 = _offsets(series_n_i_corr)
        for i in range(len(series_n_i_corr) if series_n_i_corr is not None else 0):
            for a in range(b2[i], b2[i + 1]):
                r = ele_n_i_corr[a]
                V[i] += (parameter * r[3] * (r[4] * corr_matrix[int(r[5]) - 1] + r[6] * corr_matrix[int(r[7]) - 1])
                         * np.exp(-s2 * r[2]) * _phase(gx, gy, r[0], r[1]))
    lob = paraflag == "l"
    b1 = _offsets(series_n)
    for i in range(len(series_n) if series_n is not None else 0):
        for a in range(b1[i], b1[i + 1]):
            r = ele_n[a]
            if lob:
                f = _lobato_sum(r[3:13], g2) * np.exp(-s2 * r[13])
            else:
                f = _peng_sum(r[3:13], s2)
            if absorp_n is not None:
                f = f + 1j * _peng_sum(absorp_n[a], s2)
            V[i] += parameter * r[2] * f * _phase(gx, gy, r[0], r[1])
    b2 = _offsets(series_n_i)
    for i in range(len(series_n_i) if series_n_i is not None else 0):
        for a in range(b2[i], b2[i + 1]):
            r = ele_n_i[a]
            if lob:
                f = (r[3] * _lobato_sum(r[4:14], g2) + r[14] * _lobato_sum(r[15:25], g2)) * np.exp(-s2 * r[25])
            else:
                f = r[3] * _peng_sum(r[4:14], s2) + r[14] * _peng_sum(r[15:25], s2)
            if absorp_n_i is not None:
                f = f + 1j * _peng_sum(absorp_n_i[a], s2)
            V[i] += parameter * r[2] * f * _phase(gx, gy, r[0], r[1])
    return V


def real_space_potential(Vk, aperture):
    """``..._peng.m:188``: ifft2(ifftshift(V_k .* APERTURE)) * Np^2 (unnormalised inverse DFT)."""
    n0, n1 = Vk.shape[-2:]
    return np.fft.ifft2(np.fft.ifftshift(Vk * aperture, axes=(-2, -1)), axes=(-2, -1)) * (n0 * n1)


def transmission(V, sigma):
    """``..._peng.m:189``: t = exp(i * V * sigma / 1000)."""
    return np.exp(1j * V * sigma / 1000.0)


def potential_slices(np_side, ele_n, absorp_n, ele_n_i, absorp_n_i, series_n, series_n_i,
                     s2, gx, gy, sigma, parameter, aperture, paraflag, **corr):
    """The whole ``GetPotential4AllSlice_multicore_lotabo_peng(_corr)``: returns (V, t) [S, Np, Np]."""
    Vk = reciprocal_potential(np_side, ele_n, absorp_n, ele_n_i, absorp_n_i, series_n, series_n_i,
                              s2, gx, gy, parameter, paraflag, **corr)
    V = real_space_potential(Vk, aperture)
    return V, transmission(V, sigma)
