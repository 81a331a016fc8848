"""Electron scattering-factor parameter tables (FP64).

The numbers are the published parameterisations the reference ships as MATLAB
data files, transcribed to CSV by ``tests/golden/extract_tables.py``:

* Lobato & Van Dyck (2014), ``Lobato_para.m:4-315``: f(g) = sum_i a_i (2 + b_i g^2) / (1 + b_i g^2)^2,
  g in 1/Angstrom, Z = 1..103, stored as [a1..a5, b1..b5].
* Peng et al. (1996) neutral atoms, ``ABSFpara.m:5-220``: f(s) = sum_i a_i exp(-b_i s^2), s = g/2,
  Z = 1..98, stored as [a1..a5, b1..b5].
* Peng ions, ``ABSFpara.m:224-436``: rows [Z, valence, a1..a5, b1..b5].
"""
import functools
import os

import numpy as np

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


@functools.lru_cache(maxsize=None)
def _load(name):
    return np.loadtxt(os.path.join(_DATA, name), delimiter=",", comments="#", dtype=np.float64)


def lobato_table():
    """[103, 10] array: row Z-1 holds a1..a5, b1..b5 (``Lobato_para.m:313-315``)."""
    return _load("lobato.csv")


def peng_table():
    """[98, 10] array: row Z-1 holds a1..a5, b1..b5 (``ABSFpara.m:5-220``)."""
    return _load("peng_neutral.csv")


def peng_ion_table():
    """[n, 12] array of [Z, valence, a1..a5, b1..b5] (``ABSFpara.m:224-436``)."""
    return _load("peng_ion.csv")


def interleaved(row10):
    """[a1..a5, b1..b5] -> [a1, b1, a2, b2, ..., a5, b5].

    This is the coefficient order the reference's atom matrices use
    (``getpotentialpara_lotabo_peng.m:27-30`` picks columns 1,6,2,7,...,5,10).
    """
    row10 = np.asarray(row10, dtype=np.float64)
    out = np.empty(row10.shape, dtype=np.float64)
    out[..., 0::2] = row10[..., :5]
    out[..., 1::2] = row10[..., 5:]
    return out


def lobato_f(Z, g2):
    """Lobato electron scattering factor f(g) [Angstrom] at g^2 [1/Angstrom^2]."""
    p = lobato_table()[Z - 1]
    a, b = p[:5], p[5:]
    g2 = np.asarray(g2, dtype=np.float64)[..., None]
    return np.sum(a * (2.0 + b * g2) / (1.0 + b * g2) ** 2, axis=-1)


def peng_f(Z, s2):
    """Peng electron scattering factor f(s) [Angstrom] at s^2 = g^2/4 [1/Angstrom^2]."""
    p = peng_table()[Z - 1]
    a, b = p[:5], p[5:]
    s2 = np.asarray(s2, dtype=np.float64)[..., None]
    return np.sum(a * np.exp(-b * s2), axis=-1)
