"""Frozen-phonon displacement and specimen slicing restated in FP64 (oracle; test infrastructure only).

Displacement (``randshiftpos.m:1-44``): every atom moves by a Gaussian of rms sigma per axis and is
drawn again while it leaves the slab [st[0], st[S-1] + dz) (``:7-14``) or ends on the other side of a
mid-slice face st[m-1] + dz than it started (``:16-44``).  The device draws with a counter-based
generator so that a (seed, configuration) pair names one specimen: Philox4x32-10 (Salmon, Moraes,
Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11) keyed by the seed, counter
(atom, configuration, attempt, lane); 53-bit uniforms u = ((a >> 5) 2^26 + (b >> 6) + 1/2) / 2^53;
Box-Muller normals (lane 0 gives x and y, lane 1 gives z); at most 64 attempts, after which the
atom keeps its position.  The reference redraws with MATLAB's shuffled ``randn`` -- no seed
reproduces it, so the pins here are the generator's published known-answer vectors and the
statistics of the draws.

Slicing (``CalAllEquilent20210223_lobato_peng.m:23-57`` with multiceng = 0): slice i holds the
atoms with st[i] <= z < st[i] + dz, st[i] = st[0] + dz i.

Pins (``tests/test_phonon.py``): Philox4x32-10 against the Random123 known-answer vectors; the
normals' mean, variance and independence; every accepted position obeys the slab and face
constraints; slicing against brute force on atoms placed on and next to slice faces.
"""
import math

import numpy as np

M0, M1 = 0xD2511F53, 0xCD9E8D57
W0, W1 = 0x9E3779B9, 0xBB67AE85
MASK = 0xFFFFFFFF
MAX_ATTEMPTS = 64


def philox4x32_10(ctr, key):
    """Ten Philox4x32 rounds of counter ``ctr`` (4 words) under ``key`` (2 words)."""
    c0, c1, c2, c3 = (int(v) & MASK for v in ctr)
    k0, k1 = (int(v) & MASK for v in key)
    for _ in range(10):
        p0, p1 = M0 * c0, M1 * c2
        c0, c1, c2, c3 = ((p1 >> 32) ^ c1 ^ k0) & MASK, p1 & MASK, ((p0 >> 32) ^ c3 ^ k1) & MASK, p0 & MASK
        k0, k1 = (k0 + W0) & MASK, (k1 + W1) & MASK
    return c0, c1, c2, c3


def u53(a, b):
    return ((a >> 5) * 67108864.0 + (b >> 6) + 0.5) * (1.0 / 9007199254740992.0)


def normals3(seed, atom, config, attempt):
    key = (seed & MASK, (seed >> 32) & MASK)
    a = philox4x32_10((atom, config, attempt, 0), key)
    b = philox4x32_10((atom, config, attempt, 1), key)
    r, th = math.sqrt(-2.0 * math.log(u53(a[0], a[1]))), 2.0 * math.pi * u53(a[2], a[3])
    r2, th2 = math.sqrt(-2.0 * math.log(u53(b[0], b[1]))), 2.0 * math.pi * u53(b[2], b[3])
    return r * math.cos(th), r * math.sin(th), r2 * math.cos(th2)


def displace(xyz, rms, seed, config, lo, hi, faces=()):
    """Displaced positions [n, 3] of configuration ``config``."""
    xyz = np.asarray(xyz, dtype=np.float64)
    out = xyz.copy()
    for a in range(len(xyz)):
        s = 0.0 if rms is None else float(rms[a])
        if s <= 0.0:
            continue
        x, y, z = xyz[a]
        for k in range(MAX_ATTEMPTS):
            g = normals3(seed, a, config, k)
            cx, cy, cz = x + s * g[0], y + s * g[1], z + s * g[2]
            ok = lo <= cz < hi and all((z < f) == (cz < f) for f in faces)
            if ok:
                out[a] = (cx, cy, cz)
                break
    return out


def slab(origin, dz, S, mid=()):
    """(lo, hi, faces) of S slices of thickness dz from origin, with 1-based mid-slice faces."""
    return origin, origin + dz * (S - 1) + dz, tuple(origin + dz * (m - 1) + dz for m in mid)


def slice_members(z, origin, dz, S):
    """For each slice i, the indices of z with st[i] <= z < st[i] + dz (brute force)."""
    z = np.asarray(z, dtype=np.float64)
    st = origin + dz * np.arange(S)
    return [np.nonzero((z >= st[i]) & (z < st[i] + dz))[0] for i in range(S)]
