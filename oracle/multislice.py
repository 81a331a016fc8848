"""FP64 restatement of the reference CPU multislice STEM engine (oracle; test infrastructure only).

Follows ``cuda_STEM_CPU_Mcore_lobato_peng.m:50-97`` (identical to ``..._corr.m:60-107``):

    APER := ifftshift(APER)                                    (:50-52)
    for scan column i, scan row j, focal member nn:            (:55-57)
        wave = probe[nn]
        for slice k = 1..S:
            wave = wave .* t_k(window at row j*step + r0, column i*step + c0)   (:60-61)
            wave = fft2(wave); wave = wave .* psf_fft                          (:62-63)
            if k in midslice: mid[j, i, kk, oo] += gfsf[nn] * sum |wave .* APER_kk|^2  (:65-70)
            wave = ifft2(wave)                                                  (:72)
        wave = fft2(wave)                                                       (:74)
        res[j, i, kk] += gfsf[nn] * sum |wave .* APER_kk|^2                     (:75-77)

and the output flattening ``:84-97``: result[d, a] with d = i + j*width_red, mid
[(d + a*D), l].  ``probe_row0``/``probe_col0`` are 1-based as in MATLAB.

Pins (``tests/test_oracle.py``): vacuum specimen == closed form from the probe;
brute-force 2-D DFT on an 8x8 grid; mid-slice at k == final output of a k-slice
specimen; focal-series linearity; scan/translation equivalence.
"""
import numpy as np


def stem_position(t, probe, gfsf, psf_fft, aper_shifted, r0, c0, midslice=(), periodic=False, waves=False):
    """One scan position, all focal members: (final [A], mid [L, A]) in FP64, and with
    ``waves`` also the exit waves fft2(psi) [M, n, n] the final detector integrates (``:74``).

    ``periodic`` takes the window modulo the potential side (the periodic extension
    of the reciprocal-space potential); the reference crops without wrapping."""
    S = t.shape[0]
    n = probe.shape[-1]
    if periodic:
        npx = t.shape[-1]
        ri = (r0 + np.arange(n)) % npx
        ci = (c0 + np.arange(n)) % npx
        window = lambda k: t[k - 1][np.ix_(ri, ci)]
    else:
        window = lambda k: t[k - 1, r0:r0 + n, c0:c0 + n]
    A = aper_shifted.shape[0]
    mids = list(midslice)
    res = np.zeros(A)
    mid = np.zeros((len(mids), A))
    exits = []
    for nn in range(len(gfsf)):
        wave = probe[nn].astype(np.complex128)
        for k in range(1, S + 1):
            wave = wave * window(k)
            wave = np.fft.fft2(wave)
            wave = wave * psf_fft
            if k in mids:
                oo = mids.index(k)
                for kk in range(A):
                    mid[oo, kk] += gfsf[nn] * np.sum(np.abs(wave * aper_shifted[kk]) ** 2)
            wave = np.fft.ifft2(wave)
        wave = np.fft.fft2(wave)
        exits.append(wave)
        for kk in range(A):
            res[kk] += gfsf[nn] * np.sum(np.abs(wave * aper_shifted[kk]) ** 2)
    if waves:
        return res, mid, np.asarray(exits)
    return res, mid


def stem_column(args):
    """All scan rows j of one scan column i (the unit of the reference's ``parfor``)."""
    (i, t, probe, gfsf, psf_fft, aper_shifted, probe_row0, probe_col0, hight_red, probestep, midslice) = args
    out = []
    for j in range(hight_red):
        r0 = j * probestep + probe_row0 - 1
        c0 = i * probestep + probe_col0 - 1
        out.append(stem_position(t, probe, gfsf, psf_fft, aper_shifted, r0, c0, midslice))
    return out


def stem_cpu(t, probe, gfsf, psf_fft, probe_row0, probe_col0, width_red, hight_red, probestep, APER,
             midslice=(), pool=None):
    """Whole-scan FP64 STEM.  Returns (aamyresul [D, A], aamidresul [D*A, L]).

    ``pool`` may be a ``multiprocessing`` pool; columns are mapped over it like the
    reference's ``parfor i=0:width_red-1``.
    """
    aper_shifted = np.fft.ifftshift(np.asarray(APER, dtype=np.float64), axes=(-2, -1))
    A = aper_shifted.shape[0]
    L = len(midslice)
    jobs = [(i, t, probe, gfsf, psf_fft, aper_shifted, probe_row0, probe_col0, hight_red, probestep,
             tuple(midslice)) for i in range(width_red)]
    cols = pool.map(stem_column, jobs) if pool is not None else [stem_column(j) for j in jobs]
    D = width_red * hight_red
    res = np.zeros((D, A))
    mid = np.zeros((D * A, L))
    for i, col in enumerate(cols):
        for j, (r, m) in enumerate(col):
            d = i + j * width_red
            res[d] = r
            for a in range(A):
                mid[d + a * D, :] = m[:, a]
    return res, mid
