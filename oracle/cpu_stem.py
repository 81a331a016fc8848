"""ctypes binding of the FP64 C++ oracle (``oracle/cpu_stem.cpp``) -- test infrastructure only.

The same arithmetic as :mod:`oracle.potential` and :mod:`oracle.multislice` (the
reference's MATLAB multicore CPU path) in compiled FP64 with OpenMP over slices
(potential) and probe positions (propagation), like the reference's ``parfor``
loops.  ``bench.py`` times it as the CPU baseline; ``tests/test_oracle_cpp.py``
checks it against the numpy oracle.  Built by ``make -C oracle``.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_stem.so")

_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)


class _PotDesc(ctypes.Structure):
    _fields_ = [
        ("paraflag", ctypes.c_int), ("np", ctypes.c_int), ("slices", ctypes.c_int),
        ("series_n", _DP), ("ele_n", _DP), ("absorp_n", _DP),
        ("series_n_i", _DP), ("ele_n_i", _DP), ("absorp_n_i", _DP),
        ("series_n_corr", _DP), ("ele_n_corr", _DP),
        ("series_n_i_corr", _DP), ("ele_n_i_corr", _DP),
        ("corr_matrix", _DP), ("cim", ctypes.c_int),
        ("gx", _DP), ("gy", _DP), ("s2", _DP), ("aperture", _DP),
        ("sigma", ctypes.c_double), ("parameter", ctypes.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("C++ oracle not built: %s (make -C oracle)" % LIB)
        L = ctypes.CDLL(LIB)
        L.oracle_fft2.argtypes = [_DP, ctypes.c_int, ctypes.c_int]
        L.oracle_potential.argtypes = [ctypes.POINTER(_PotDesc), ctypes.c_int, ctypes.c_int, _DP, _DP, ctypes.c_int]
        L.oracle_stem.argtypes = [_DP, ctypes.c_int, ctypes.c_int, _DP, ctypes.c_int, _DP, _DP, _DP, ctypes.c_int,
                                  ctypes.c_int, _IP, _IP, ctypes.c_int, ctypes.c_int, _IP, ctypes.c_int, _DP, _DP,
                                  ctypes.c_int]
        L.oracle_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _f64(a):
    return None if a is None else np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _c128(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.complex128))


def _dp(a):
    return None if a is None else a.ctypes.data_as(_DP)


def fft2(x, sign):
    """Unnormalised 2-D DFT of a square complex matrix (sign -1 forward, +1 inverse)."""
    y = _c128(x).copy()
    rc = lib().oracle_fft2(y.ctypes.data_as(_DP), y.shape[0], int(sign))
    if rc:
        raise ValueError("oracle_fft2: unsupported size %d" % y.shape[0])
    return y


def potential_slices(np_side, ele_n, absorp_n, ele_n_i, absorp_n_i, series_n, series_n_i, s2, gx, gy, sigma,
                     parameter, aperture, paraflag, ele_n_corr=None, ele_n_i_corr=None, corr_matrix=None,
                     series_n_corr=None, series_n_i_corr=None, slices=None, threads=None, want_v=True):
    """Same arguments and result as :func:`oracle.potential.potential_slices`: (V, t) [S, Np, Np]
    complex128.  ``slices`` = (s0, s1) builds only that range (the rest stays zero)."""
    keep = []

    def p(a):
        a = _f64(a)
        keep.append(a)
        return _dp(a)

    S = max(len(series_n) if series_n is not None else 0, len(series_n_i) if series_n_i is not None else 0)
    d = _PotDesc()
    d.paraflag, d.np, d.slices = ord(paraflag), int(np_side), S
    d.series_n, d.ele_n, d.absorp_n = p(series_n), p(ele_n), p(absorp_n)
    d.series_n_i = p(series_n_i if series_n_i is not None and len(series_n_i) else None)
    d.ele_n_i = p(ele_n_i if ele_n_i is not None and len(ele_n_i) else None)
    d.absorp_n_i = p(absorp_n_i)
    corr = paraflag == "n" and corr_matrix is not None and len(corr_matrix)
    d.series_n_corr = p(series_n_corr if corr else None)
    d.ele_n_corr = p(ele_n_corr if corr and ele_n_corr is not None and len(ele_n_corr) else None)
    d.series_n_i_corr = p(series_n_i_corr if corr else None)
    d.ele_n_i_corr = p(ele_n_i_corr if corr and ele_n_i_corr is not None and len(ele_n_i_corr) else None)
    d.corr_matrix = p(corr_matrix if corr else None)
    d.cim = int(np.asarray(corr_matrix).shape[0]) if corr else 0
    d.gx, d.gy, d.s2, d.aperture = p(gx), p(gy), p(s2), p(aperture)
    d.sigma, d.parameter = float(sigma), float(parameter)
    s0, s1 = (0, S) if slices is None else slices
    t = np.zeros((S, np_side, np_side), dtype=np.complex128)
    v = np.zeros_like(t) if want_v else None
    rc = lib().oracle_potential(ctypes.byref(d), int(s0), int(s1), t.ctypes.data_as(_DP),
                                None if v is None else v.ctypes.data_as(_DP),
                                int(threads or lib().oracle_max_threads()))
    if rc:
        raise ValueError("oracle_potential failed (%d)" % rc)
    return v, t


def stem_positions(t, probe, gfsf, psf_fft, aper_shifted, r0, c0, midslice=(), periodic=False, threads=None):
    """:func:`oracle.multislice.stem_position` for many positions at once (0-based window
    origins r0[d], c0[d]): (final [D, A], mid [D, L, A])."""
    t = _c128(t)
    probe = _c128(probe)
    gfsf = _f64(gfsf)
    psf = _c128(psf_fft)
    ap = _f64(aper_shifted)
    r0 = np.ascontiguousarray(np.asarray(r0, dtype=np.int32))
    c0 = np.ascontiguousarray(np.asarray(c0, dtype=np.int32))
    mid = np.ascontiguousarray(np.asarray(midslice, dtype=np.int32))
    S, npx = t.shape[0], t.shape[-1]
    M, n = probe.shape[0], probe.shape[-1]
    A, D, L = ap.shape[0], len(r0), len(mid)
    out = np.zeros((D, A))
    mid_out = np.zeros((D, max(L, 1), A))
    rc = lib().oracle_stem(t.ctypes.data_as(_DP), npx, S, probe.ctypes.data_as(_DP), M, _dp(gfsf),
                           psf.ctypes.data_as(_DP), _dp(ap), A, n, r0.ctypes.data_as(_IP), c0.ctypes.data_as(_IP),
                           D, int(bool(periodic)), mid.ctypes.data_as(_IP) if L else None, L, _dp(out),
                           _dp(mid_out), int(threads or lib().oracle_max_threads()))
    if rc:
        raise ValueError("oracle_stem failed (%d)" % rc)
    return out, mid_out[:, :L]


def max_threads():
    return int(lib().oracle_max_threads())
