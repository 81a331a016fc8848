"""ctypes binding of the engine's C ABI (include/totem_abi.h).

Mirrors the reference's mex call shape:

    [myresul, mid_ceng_mat, potentialx, potentialy] = waveCudalobato_<v>(41 args)

(``waveCuda.cpp:4-101``, driven from ``cuda_stem.m:81-107``) as
:func:`wave_cuda`, with numpy arrays in the row-major layout that
``cuda_stem.m:22-79`` hands to the mex.  The library is the in-tree
``totem_amd/lib/libtotem.so``; importing this module on a machine without it
raises immediately (there is no CPU fallback).
"""
import ctypes
import os

import numpy as np

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
CORE = os.environ.get("TOTEM_LIB") or os.path.join(LIB_DIR, "libtotem.so")
VARIANT_LIBS = {
    "waveCudalobato": "libwaveCudalobato.so",
    "waveCudaPengAndCorr": "libwaveCudaPengAndCorr.so",
    "waveCudalobatoP": "libwaveCudalobatoP.so",
    "waveCudaPengAndCorrP": "libwaveCudaPengAndCorrP.so",
}
ENTRY = {
    "waveCudalobato": "totem_project_lobato",
    "waveCudaPengAndCorr": "totem_project_peng",
    "waveCudalobatoP": "totem_potential_lobato",
    "waveCudaPengAndCorrP": "totem_potential_peng",
}
# the symbols include/totem_abi.h declares
EXPORTS = (
    "myProjectfun",
    "totem_project_lobato", "totem_project_peng", "totem_potential_lobato", "totem_potential_peng",
    "totem_last_error", "totem_abi_version", "totem_fft_size_supported", "totem_build_transmission",
    "totem_potential_plan_stats",
    "totem_stem_workspace_bytes", "totem_stem_run", "totem_stem_run_static", "totem_kernel_times", "totem_fft2d",
    "totem_stem_preferred_batch", "totem_stem_autotune", "totem_stem_plan_count", "totem_stem_plan",
    "totem_stem_plan_columns", "totem_stem_band",
    "totem_stem_set_plan", "totem_hrtem_images", "totem_tcc_image", "totem_weighted_pattern",
    "totem_server_create", "totem_server_submit", "totem_server_step", "totem_server_poll", "totem_server_sync",
    "totem_server_destroy",
    "totem_specimen_create", "totem_specimen_transmission", "totem_specimen_positions", "totem_specimen_destroy",
)

TOTEM_RUN_GRAPH = 1
TOTEM_RUN_TIMING = 2
ABI_VERSION = 2   # layout of the structs below (totem_abi_version())


class TotemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("totem error %d: %s" % (code, msg))
        self.code = code


_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)
_VP = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double

# the 43 parameters of myProjectfun, in prototype order
_PROJECT_ARGTYPES = [_DP] * 15 + [_I] * 10 + [_D, _I, _I, _D, _DP, _IP, _I] + [_DP] * 6 + [_I] + [_DP] * 4


class PotentialDesc(ctypes.Structure):
    _fields_ = [
        ("paraflag", _I), ("np", _I), ("slices", _I),
        ("atom_slice", _DP), ("proj_coff_mat", _DP), ("absorp_n", _DP),
        ("atom_slice_i", _DP), ("proj_coff_mat_i", _DP), ("absorp_n_i", _DP),
        ("series_n_corr", _DP), ("ele_n_corr", _DP), ("series_n_i_corr", _DP), ("ele_n_i_corr", _DP),
        ("corr_info_matrix", _DP), ("cim_len", _I),
        ("kx", _DP), ("ky", _DP), ("s_2", _DP), ("aper_true", _DP),
        ("sigma", _D), ("parameter", _D),
    ]


class SpecimenDesc(ctypes.Structure):
    _fields_ = [
        ("pot", PotentialDesc), ("n_atoms", _I), ("xyz", _DP), ("rms", _DP), ("term_atom", _IP), ("n_terms", _I),
        ("slices", _I), ("slice_origin", _D), ("slice_thickness", _D), ("mid", _IP), ("n_mid", _I),
        ("multiceng", _D),
    ]


class StemJob(ctypes.Structure):
    _fields_ = [
        ("n", _I), ("np", _I), ("slices", _I),
        ("t", _VP), ("probes", _VP), ("n_probes", _I), ("prop", _VP),
        ("aper", _VP), ("aper2", _VP), ("n_aper", _I),
        ("mid", _IP), ("n_mid", _I),
        ("depth", _I), ("width_red", _I), ("step", _I), ("begin_row", _I), ("begin_col", _I), ("wrap", _I),
        ("det", _VP), ("exit_wave", _VP),
    ]


class HrtemDesc(ctypes.Structure):
    _fields_ = [
        ("n", _I), ("batches", _I), ("mtotal", _I),
        ("lambda_nm", ctypes.c_double), ("alafa", ctypes.c_double), ("delta_yita", ctypes.c_double),
        ("gfsf", ctypes.c_double),
        ("u", _DP), ("v", _DP), ("wave", _DP), ("pp2", _DP), ("mrr", _DP), ("ai", _DP),
    ]


_lib = None


def lib():
    """Load libtotem.so (in-tree build).  Raises if it is missing."""
    global _lib
    if _lib is None:
        try:
            # torch ships its own HIP runtime; importing it first makes the engine library bind to
            # that same runtime (one runtime per process, shared device pointers and streams)
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(CORE):
            raise RuntimeError("totem engine library not built: %s (run __graft_entry__.build())" % CORE)
        L = ctypes.CDLL(CORE, mode=ctypes.RTLD_GLOBAL)
        for name in ENTRY.values():
            getattr(L, name).argtypes = _PROJECT_ARGTYPES
            getattr(L, name).restype = _I
        L.totem_last_error.restype = ctypes.c_char_p
        L.totem_abi_version.restype = _I
        L.totem_fft_size_supported.argtypes = [_I]
        L.totem_fft_size_supported.restype = _I
        L.totem_build_transmission.argtypes = [ctypes.POINTER(PotentialDesc), _VP, _DP, _DP, _VP]
        L.totem_build_transmission.restype = _I
        L.totem_stem_workspace_bytes.argtypes = [ctypes.POINTER(StemJob), _I]
        L.totem_stem_workspace_bytes.restype = ctypes.c_size_t
        L.totem_stem_run.argtypes = [ctypes.POINTER(StemJob), _I, _I, _VP, ctypes.c_size_t, _VP]
        L.totem_stem_run.restype = _I
        L.totem_stem_run_static.argtypes = [ctypes.POINTER(StemJob), _I, _VP]
        L.totem_stem_run_static.restype = _I
        L.totem_kernel_times.argtypes = [_I, _DP, ctypes.POINTER(ctypes.c_long), _DP, ctypes.POINTER(ctypes.c_long)]
        L.totem_kernel_times.restype = None
        L.totem_fft2d.argtypes = [_VP, _I, _I, _I, _VP]
        L.totem_fft2d.restype = _I
        L.totem_hrtem_images.argtypes = [ctypes.POINTER(HrtemDesc), _DP, _DP, _VP]
        L.totem_hrtem_images.restype = _I
        L.totem_tcc_image.argtypes = [_I, _I, _DP, _DP, _DP, _DP, _DP, _VP]
        L.totem_tcc_image.restype = _I
        L.totem_weighted_pattern.argtypes = [_I, _I, _VP, _DP, _I, _DP, _VP]
        L.totem_weighted_pattern.restype = _I
        L.totem_stem_autotune.argtypes = [_I, _I, ctypes.POINTER(_I), _DP]
        L.totem_stem_autotune.restype = _I
        L.totem_stem_preferred_batch.argtypes = [_I]
        L.totem_stem_preferred_batch.restype = _I
        L.totem_stem_plan_count.argtypes = [_I]
        L.totem_stem_plan_count.restype = _I
        L.totem_stem_plan.argtypes = [_I, ctypes.POINTER(_I), ctypes.c_char_p, ctypes.c_size_t]
        L.totem_stem_plan.restype = _I
        L.totem_stem_set_plan.argtypes = [_I, _I]
        L.totem_stem_plan_columns.argtypes = [_I, _I]
        L.totem_stem_plan_columns.restype = _I
        L.totem_stem_band.argtypes = [ctypes.POINTER(_I), _I, _I, ctypes.POINTER(_I), ctypes.POINTER(_I)]
        L.totem_stem_set_plan.restype = _I
        L.totem_specimen_create.argtypes = [ctypes.POINTER(SpecimenDesc), ctypes.POINTER(_VP)]
        L.totem_specimen_transmission.argtypes = [_VP, ctypes.c_ulonglong, _I, _VP, _VP]
        L.totem_specimen_positions.argtypes = [_VP, ctypes.c_ulonglong, _I, _DP]
        L.totem_specimen_destroy.argtypes = [_VP]
        for name in ("create", "transmission", "positions", "destroy"):
            getattr(L, "totem_specimen_" + name).restype = _I
        L.totem_server_create.argtypes = [ctypes.c_size_t, _I, _I, ctypes.POINTER(_VP)]
        L.totem_server_submit.argtypes = [_VP, ctypes.POINTER(StemJob), ctypes.POINTER(ctypes.c_longlong)]
        L.totem_server_step.argtypes = [_VP, ctypes.POINTER(_I)]
        L.totem_server_poll.argtypes = [_VP, ctypes.POINTER(ctypes.c_longlong), _I, ctypes.POINTER(_I)]
        L.totem_server_sync.argtypes = [_VP]
        L.totem_server_destroy.argtypes = [_VP]
        for name in ("create", "submit", "step", "poll", "sync", "destroy"):
            getattr(L, "totem_server_" + name).restype = _I
        if L.totem_abi_version() != ABI_VERSION:
            raise RuntimeError("%s has ABI version %d, this binding expects %d (rebuild)"
                               % (CORE, L.totem_abi_version(), ABI_VERSION))
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise TotemError(rc, lib().totem_last_error().decode())


def _dp(a):
    """double* view of a C-contiguous float64 array (or NULL)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(_DP)


def _f64(a):
    if a is None:
        return None
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


_drop_in = {}


def drop_in_library(name):
    """ctypes handle of the drop-in library ``lib<name>.so`` (``name`` a key of ENTRY), whose one
    export is ``myProjectfun`` with the reference's prototype (``waveCuda.h:10-21``)."""
    h = _drop_in.get(name)
    if h is None:
        lib()  # the engine first: the drop-in library then binds to this same libtotem.so
        h = ctypes.CDLL(os.path.join(os.path.dirname(CORE), "lib%s.so" % name))
        h.myProjectfun.argtypes = _PROJECT_ARGTYPES
        h.myProjectfun.restype = None
        _drop_in[name] = h
    return h


def my_project_fun(entry, my_real_w, my_imag_w, aper, atom_slice, absorp_n, atom_slice_i, absorp_n_i,
                   proj_coff_mat, proj_coff_mat_i, kx, ky, s_2, my_real_const, my_img_const, gfsf,
                   Height, Width, Depth, P_Height, P_Width, Slices, step, beginRow, beginCol, width_red,
                   sigma, w_num, aper_num, parameter, AperTrue, mid_ceng, ceng_len, aper2,
                   series_n_corr, series_n_i_corr, ele_n_corr, ele_n_i_corr, corr_info_matrix, cim_len,
                   drop_in=None):
    """Call one engine entry point with the 39 inputs of myProjectfun (``waveCuda.h:10-21``);
    beginRow/beginCol are 0-based here, as in the C prototype.  With ``drop_in`` (a key of ENTRY)
    the call goes through that drop-in library's ``myProjectfun`` symbol instead, which returns
    nothing and leaves a failure in totem_last_error().  Returns
    (return_result [aper_num, Depth], mid_ceng_mat [ceng_len, aper_num, Depth],
    potentialx, potentialy [Slices, P_Height, P_Width])."""
    arrays = [_f64(v) for v in (my_real_w, my_imag_w, aper, atom_slice, absorp_n, atom_slice_i, absorp_n_i,
                                proj_coff_mat, proj_coff_mat_i, kx, ky, s_2, my_real_const, my_img_const, gfsf)]
    ap_true = _f64(AperTrue)
    mid = np.ascontiguousarray(np.asarray(mid_ceng if mid_ceng is not None else [], dtype=np.int32))
    tail = [_f64(v) for v in (aper2, series_n_corr, series_n_i_corr, ele_n_corr, ele_n_i_corr, corr_info_matrix)]
    res = np.zeros((aper_num, Depth))
    midm = np.zeros((max(ceng_len, 0), aper_num, Depth))
    potential_only = entry.startswith("totem_potential")
    np2 = P_Height * P_Width
    px = np.zeros((Slices, P_Height, P_Width)) if potential_only else np.zeros(1)
    py = np.zeros((Slices, P_Height, P_Width)) if potential_only else np.zeros(1)
    fn = drop_in_library(drop_in).myProjectfun if drop_in else getattr(lib(), entry)
    rc = fn(*[_dp(a) for a in arrays], int(Height), int(Width), int(Depth), int(P_Height), int(P_Width),
            int(Slices), int(step), int(beginRow), int(beginCol), int(width_red), float(sigma), int(w_num),
            int(aper_num), float(parameter), _dp(ap_true), mid.ctypes.data_as(_IP), int(ceng_len),
            *[_dp(a) for a in tail], int(cim_len), _dp(res), _dp(midm) if midm.size else _dp(np.zeros(1)),
            _dp(px), _dp(py))
    if drop_in:
        msg = lib().totem_last_error().decode()
        if msg:
            raise TotemError(-1, msg)
    else:
        check(rc)
    del np2
    return res, midm, (px if potential_only else None), (py if potential_only else None)


def wave_cuda(name, my_real_w, my_img_w, aper, atom_slice, absorp_n2, atom_slice_i, absorp_n_i2, proj_coff_mat,
              proj_coff_mat_i, Kx, Ky, S_2, my_real_const, my_img_const, gfsf, Height, Width, Depth, P_Height,
              P_Width, Slices, Step, beginRow, beginCol, width_red, sigma, w_num, aper_num, parameter, aperTrue,
              absorp_flag, absorp_i_flag, mid_ceng, ceng_len, aper2, seriesn_corr, seriesn_i_corr, elen_corr,
              elen_i_corr, corrinfo_matrix, cimlen, drop_in=False):
    """The mex function ``waveCuda*_<v>`` with its 41 arguments in ``cuda_stem.m:97`` order
    (``waveCuda.cpp:18-79``: absorption flags select NULL, cimlen==0 drops the correction table,
    1-based beginRow/beginCol).  Returns the four mex outputs shaped as MATLAB sees them:
    myresul [Depth, aper_num], mid_ceng_mat [Depth*aper_num, ceng_len],
    potentialx/potentialy [P_Height*P_Width, Slices] (zeros for the STEM variants).
    With ``drop_in`` the call goes through ``myProjectfun`` of ``lib<name>.so``, as the mex
    gateway does."""
    if name not in ENTRY:
        raise ValueError("unknown mex name %r" % name)
    res, midm, px, py = my_project_fun(
        ENTRY[name], my_real_w, my_img_w, aper, atom_slice, absorp_n2 if absorp_flag == 1 else None,
        atom_slice_i, absorp_n_i2 if absorp_i_flag == 1 else None, proj_coff_mat, proj_coff_mat_i, Kx, Ky, S_2,
        my_real_const, my_img_const, gfsf, Height, Width, Depth, P_Height, P_Width, Slices, Step, beginRow - 1,
        beginCol - 1, width_red, sigma, w_num, aper_num, parameter, aperTrue, mid_ceng, ceng_len, aper2,
        seriesn_corr, seriesn_i_corr, elen_corr, elen_i_corr, corrinfo_matrix if cimlen != 0 else None, cimlen,
        drop_in=name if drop_in else None)
    L = max(int(ceng_len), 0)
    myresul = res.T.copy()                                   # [Depth, aper_num]
    mid_mat = midm.reshape(L, aper_num * Depth).T.copy()     # [Depth*aper_num, ceng_len]
    np2 = P_Height * P_Width
    if px is None:
        px = np.zeros((np2, Slices))
        py = np.zeros((np2, Slices))
    else:
        px = px.reshape(Slices, np2).T.copy()
        py = py.reshape(Slices, np2).T.copy()
    return myresul, mid_mat, px, py


# ------------------------------------------------------------------ device API


def build_transmission(paraflag, np_side, slices, atom_slice, proj_coff_mat, kx, ky, s_2, aper_true, sigma,
                       parameter, t_dev_ptr, absorp_n=None, atom_slice_i=None, proj_coff_mat_i=None,
                       absorp_n_i=None, series_n_corr=None, ele_n_corr=None, series_n_i_corr=None,
                       ele_n_i_corr=None, corr_info_matrix=None, want_v=False, stream=None):
    """Build the transmission stack t (complex64 [S, Np, Np]) into device memory at t_dev_ptr.
    Returns (V_re, V_im) host arrays when want_v, else None."""
    d, keep = _potential_desc(paraflag, np_side, slices, atom_slice, proj_coff_mat, kx, ky, s_2, aper_true, sigma,
                              parameter, absorp_n, atom_slice_i, proj_coff_mat_i, absorp_n_i, series_n_corr,
                              ele_n_corr, series_n_i_corr, ele_n_i_corr, corr_info_matrix)
    vre = vim = None
    if want_v:
        vre = np.zeros((slices, np_side, np_side))
        vim = np.zeros((slices, np_side, np_side))
    check(lib().totem_build_transmission(ctypes.byref(d), _VP(t_dev_ptr), _dp(vre), _dp(vim), _VP(stream)))
    return (vre, vim) if want_v else None


def potential_plan_stats(paraflag, np_side, slices, atom_slice, proj_coff_mat, kx, ky, s_2, aper_true, sigma,
                         parameter, **kw):
    """Host-side plan of build_transmission's arguments (no device work): a dict with the
    form-factor ``classes``, the ``terms`` over all slices, ``split`` (Peng rows split into
    one class per Gaussian term) and ``paired`` (kx, ky pair up, so a quarter of the
    reciprocal-space tiles is accumulated and mirrored)."""
    d, keep = _potential_desc(paraflag, np_side, slices, atom_slice, proj_coff_mat, kx, ky, s_2, aper_true, sigma,
                              parameter, **kw)
    classes, split, paired = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    terms = ctypes.c_longlong(0)
    check(lib().totem_potential_plan_stats(ctypes.byref(d), ctypes.byref(classes), ctypes.byref(terms),
                                           ctypes.byref(split), ctypes.byref(paired)))
    return {"classes": classes.value, "terms": terms.value, "split": bool(split.value),
            "paired": bool(paired.value)}


def _potential_desc(paraflag, np_side, slices, atom_slice, proj_coff_mat, kx, ky, s_2, aper_true, sigma, parameter,
                    absorp_n=None, atom_slice_i=None, proj_coff_mat_i=None, absorp_n_i=None, series_n_corr=None,
                    ele_n_corr=None, series_n_i_corr=None, ele_n_i_corr=None, corr_info_matrix=None):
    """A PotentialDesc and the host arrays it points into (keep them alive while it is used)."""
    keep = {}

    def p(name, a):
        a = _f64(a)
        keep[name] = a
        return _dp(a)

    d = PotentialDesc()
    d.paraflag = ord(paraflag)
    d.np = int(np_side)
    d.slices = int(slices)
    d.atom_slice = p("as", atom_slice)
    d.proj_coff_mat = p("pcm", proj_coff_mat)
    d.absorp_n = p("an", absorp_n)
    d.atom_slice_i = p("asi", atom_slice_i)
    d.proj_coff_mat_i = p("pcmi", proj_coff_mat_i)
    d.absorp_n_i = p("ani", absorp_n_i)
    d.series_n_corr = p("snc", series_n_corr)
    d.ele_n_corr = p("enc", ele_n_corr)
    d.series_n_i_corr = p("snic", series_n_i_corr)
    d.ele_n_i_corr = p("enic", ele_n_i_corr)
    d.corr_info_matrix = p("cim", corr_info_matrix)
    d.cim_len = 0 if corr_info_matrix is None else int(np.asarray(corr_info_matrix).shape[0])
    d.kx = p("kx", kx)
    d.ky = p("ky", ky)
    d.s_2 = p("s2", s_2)
    d.aper_true = p("at", aper_true)
    d.sigma = float(sigma)
    d.parameter = float(parameter)
    return d, keep


def stem_job(n, np_side, slices, t_ptr, probes_ptr, n_probes, prop_ptr, aper_ptr, n_aper, depth, width_red,
             step, begin_row, begin_col, det_ptr, aper2_ptr=None, mid=(), wrap=False, exit_ptr=None):
    """Fill a StemJob (device pointers are integers, e.g. torch ``tensor.data_ptr()``)."""
    j = StemJob()
    j.n, j.np, j.slices = int(n), int(np_side), int(slices)
    j.t, j.probes, j.n_probes, j.prop = t_ptr, probes_ptr, int(n_probes), prop_ptr
    j.aper, j.aper2, j.n_aper = aper_ptr, aper2_ptr, int(n_aper)
    mid = np.ascontiguousarray(np.asarray(mid, dtype=np.int32))
    j._mid_keep = mid
    j.mid = mid.ctypes.data_as(_IP) if mid.size else None
    j.n_mid = int(mid.size)
    j.depth, j.width_red, j.step = int(depth), int(width_red), int(step)
    j.begin_row, j.begin_col, j.wrap = int(begin_row), int(begin_col), int(bool(wrap))
    j.det = det_ptr
    j.exit_wave = exit_ptr
    return j


def preferred_batch(n):
    """Probes in flight whose n x n waves fit the GPU's memory-side cache."""
    return int(lib().totem_stem_preferred_batch(int(n)))


def hrtem_images(n, lambda_nm, alafa, delta_yita, gfsf, mtotal, U, V, mywave, pp2, mrr, ai):
    """Batch HRTEM images (B = len(ai)) and the drawn parameter rows: (images [B, N, N], paras [B, 49])."""
    ai = np.ascontiguousarray(np.atleast_2d(ai), dtype=np.float64)
    B = ai.shape[0]
    wave = np.ascontiguousarray(np.asarray(mywave, dtype=np.complex128)).view(np.float64)
    keep = [_f64(U), _f64(V), wave, _f64(pp2), _f64(mrr), ai]
    assert keep[0].shape == (n, n) and keep[1].shape == (n, n) and wave.shape == (n, 2 * n)
    assert keep[3].size == 49 and keep[4].size >= 21 and ai.shape[1] == 21
    d = HrtemDesc()
    d.n, d.batches, d.mtotal = int(n), int(B), int(mtotal)
    d.lambda_nm, d.alafa, d.delta_yita, d.gfsf = float(lambda_nm), float(alafa), float(delta_yita), float(gfsf)
    d.u, d.v, d.wave, d.pp2, d.ai = _dp(keep[0]), _dp(keep[1]), _dp(wave), _dp(keep[3]), _dp(ai)
    mrr = np.ascontiguousarray(keep[4][:21])
    keep.append(mrr)
    d.mrr = _dp(mrr)
    images = _pinned_host((B, n, n))
    paras = np.zeros((B, 49))
    check(lib().totem_hrtem_images(ctypes.byref(d), _dp(paras), _dp(images), None))
    return images, paras


def _pinned_host(shape):
    """A float64 host array in page-locked memory (torch's caching host allocator), so that large
    device-to-host copies run at the link rate instead of through pageable staging; ordinary
    pageable memory when torch is not importable."""
    try:
        import torch
    except ImportError:
        return np.zeros(shape)
    return torch.empty(shape, dtype=torch.float64, pin_memory=True).numpy()


def tcc_image(wave, tcc, ap, gfsf):
    """sum_m gfsf[m] |IFFT2(ifftshift(wave * tcc[m] * ap))|^2 (host arrays, centred inputs)."""
    wave = np.ascontiguousarray(np.asarray(wave, dtype=np.complex128))
    n = wave.shape[0]
    tcc = np.ascontiguousarray(np.asarray(tcc, dtype=np.complex128).reshape(-1, n, n))
    ap = _f64(ap)
    gfsf = _f64(gfsf)
    assert wave.shape == (n, n) and ap.shape == (n, n) and gfsf.size == tcc.shape[0]
    out = np.zeros((n, n))
    check(lib().totem_tcc_image(n, tcc.shape[0], _dp(wave.view(np.float64)), _dp(tcc.reshape(-1).view(np.float64)),
                                _dp(ap), _dp(gfsf), _dp(out), None))
    return out


def weighted_pattern(fields_ptr, n, weights, shift=True):
    """sum_m weights[m] |fields[m]|^2 of device complex64 fields [m, n, n] (fftshifted when shift)."""
    w = _f64(weights)
    out = np.zeros((n, n))
    check(lib().totem_weighted_pattern(int(n), int(w.size), _VP(fields_ptr), _dp(w), int(bool(shift)), _dp(out), None))
    return out


def autotune(n, slices=8):
    """Measure every pass plan of n x n grids with probes-per-launch around the cache-fit rule and
    adopt the fastest pair on the current device; returns (batch, probe.slice/s) (see :func:`plan`)."""
    b = _I()
    r = ctypes.c_double()
    check(lib().totem_stem_autotune(int(n), int(slices), ctypes.byref(b), ctypes.byref(r)))
    return b.value, r.value


def plan_count(n):
    """Pass plans (schedules of the rows and columns passes) compiled for n x n grids."""
    return int(lib().totem_stem_plan_count(int(n)))


def plan(n):
    """(index, description) of the pass plan in use for n x n grids on the current device."""
    p = _I()
    buf = ctypes.create_string_buffer(512)
    check(lib().totem_stem_plan(int(n), ctypes.byref(p), buf, ctypes.c_size_t(len(buf))))
    return p.value, buf.value.decode()


def set_plan(n, index):
    """Use pass plan ``index`` for n x n grids on the current device (runs started afterwards)."""
    check(lib().totem_stem_set_plan(int(n), int(index)))


def plan_columns(n, index):
    """Columns per columns-pass workgroup of pass plan ``index`` of n x n grids."""
    return int(lib().totem_stem_plan_columns(int(n), int(index)))


def stem_band(prop, group):
    """(lo, len): the wrapped run of columns [lo, lo + len) (mod n) a run of the propagator
    ``prop`` [n, n] processes with ``group`` columns per columns-pass workgroup (the columns
    holding a nonzero entry, widened to whole groups and to the shortest wrapped run covering
    them)."""
    prop = np.asarray(prop)
    n = prop.shape[-1]
    flags = np.ascontiguousarray((prop != 0).any(axis=0).astype(np.int32))
    lo, ln = _I(), _I()
    check(lib().totem_stem_band(flags.ctypes.data_as(ctypes.POINTER(_I)), int(n), int(group), ctypes.byref(lo),
                                ctypes.byref(ln)))
    return lo.value, ln.value


def stem_workspace_bytes(job, batch):
    return int(lib().totem_stem_workspace_bytes(ctypes.byref(job), int(batch)))


def stem_run(job, batch, graph=False, workspace_ptr=None, workspace_bytes=0, stream=None, timing=False):
    flags = (TOTEM_RUN_GRAPH if graph else 0) | (TOTEM_RUN_TIMING if timing else 0)
    check(lib().totem_stem_run(ctypes.byref(job), int(batch), flags, _VP(workspace_ptr),
                               ctypes.c_size_t(workspace_bytes), _VP(stream)))


STATIC_SCRATCH_BYTES = 128 << 20   # TOTEM_STATIC_SCRATCH_BYTES


def stem_run_static(job, batch, stream=None):
    """totem_stem_run_static: the run in the library's static device scratch, no allocation;
    returns when the run has finished."""
    check(lib().totem_stem_run_static(ctypes.byref(job), int(batch), _VP(stream)))


def kernel_times(reset=True):
    """(rows_ms, rows_launches, cols_ms, cols_launches) accumulated by timed runs."""
    rm, cm = ctypes.c_double(), ctypes.c_double()
    rn, cn = ctypes.c_long(), ctypes.c_long()
    lib().totem_kernel_times(int(reset), ctypes.byref(rm), ctypes.byref(rn), ctypes.byref(cm), ctypes.byref(cn))
    return rm.value, rn.value, cm.value, cn.value


def fft2d(ptr, n, batch, sign, stream=None):
    check(lib().totem_fft2d(_VP(ptr), int(n), int(batch), int(sign), _VP(stream)))


def fft_size_supported(n):
    return bool(lib().totem_fft_size_supported(int(n)))


# ------------------------------------------------------------------ served engine


class Server:
    """Continuous batching of device STEM jobs (``totem_server_*``): submit jobs at any time,
    call :meth:`step` to advance every in-flight probe by one slice, :meth:`poll` for the ids
    of finished jobs.  A job's device arrays must stay alive until it is reported finished."""

    def __init__(self, arena_bytes, max_aper=8, max_mid=16):
        h = _VP()
        check(lib().totem_server_create(ctypes.c_size_t(int(arena_bytes)), int(max_aper), int(max_mid),
                                        ctypes.byref(h)))
        self._h = h
        self._jobs = {}
        self._ids = (ctypes.c_longlong * 256)()

    def submit(self, job):
        jid = ctypes.c_longlong()
        check(lib().totem_server_submit(self._h, ctypes.byref(job), ctypes.byref(jid)))
        self._jobs[jid.value] = job   # keeps the job's host-side arrays (mid list) alive
        return jid.value

    def step(self):
        n = _I()
        check(lib().totem_server_step(self._h, ctypes.byref(n)))
        return n.value

    def poll(self):
        out = []
        while True:
            n = _I()
            check(lib().totem_server_poll(self._h, self._ids, len(self._ids), ctypes.byref(n)))
            out.extend(self._ids[k] for k in range(n.value))
            if n.value < len(self._ids):
                break
        for jid in out:
            self._jobs.pop(jid, None)
        return out

    @property
    def busy(self):
        return bool(self._jobs)

    def sync(self):
        check(lib().totem_server_sync(self._h))

    def close(self):
        if self._h:
            check(lib().totem_server_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceSpecimen:
    """A specimen resident on the GPU that is displaced (frozen phonon) and sliced there per
    configuration (``totem_specimen_*``).  ``rows`` holds every atom's coefficient rows as one slice
    (``stem.specimen_rows``), ``term_atom`` the atom of each row (ele_n, ele_n_i, ele_n_corr order)."""

    def __init__(self, paraflag, np_side, rows, term_atom, xyz, rms, kx, ky, s_2, aper_true, sigma, parameter,
                 slices, slice_origin, slice_thickness, mid=(), corr_info_matrix=None, multiceng=0.0):
        d, keep = _potential_desc(paraflag, np_side, 1, rows["series_n"], rows["ele_n"], kx, ky, s_2, aper_true,
                                  sigma, parameter, atom_slice_i=rows["series_n_i"], proj_coff_mat_i=rows["ele_n_i"],
                                  series_n_corr=rows.get("series_n_corr"), ele_n_corr=rows.get("ele_n_corr"),
                                  corr_info_matrix=corr_info_matrix)
        sd = SpecimenDesc()
        sd.pot = d
        xyz = _f64(xyz).reshape(-1, 3)
        keep["xyz"] = xyz
        sd.n_atoms = xyz.shape[0]
        sd.xyz = _dp(xyz)
        if rms is not None:
            rms = _f64(rms).reshape(-1)
            if rms.shape[0] != sd.n_atoms:
                raise ValueError("rms needs one entry per atom")
        keep["rms"] = rms
        sd.rms = _dp(rms)
        ta = np.ascontiguousarray(np.asarray(term_atom, dtype=np.int32))
        keep["ta"] = ta
        sd.term_atom = ta.ctypes.data_as(_IP) if ta.size else None
        sd.n_terms = int(ta.size)
        sd.slices = int(slices)
        sd.slice_origin = float(slice_origin)
        sd.slice_thickness = float(slice_thickness)
        mid = np.ascontiguousarray(np.asarray(mid, dtype=np.int32))
        keep["mid"] = mid
        sd.mid = mid.ctypes.data_as(_IP) if mid.size else None
        sd.n_mid = int(mid.size)
        sd.multiceng = float(multiceng)
        h = _VP()
        check(lib().totem_specimen_create(ctypes.byref(sd), ctypes.byref(h)))
        self._h = h
        self.n_atoms = sd.n_atoms
        self.np_side = int(np_side)
        self.slices = int(slices)

    def transmission(self, seed, config, t_dev_ptr, stream=None):
        """t (complex64 [S, Np, Np]) of configuration ``config`` into device memory at t_dev_ptr."""
        check(lib().totem_specimen_transmission(self._h, ctypes.c_ulonglong(seed), int(config), _VP(t_dev_ptr),
                                                _VP(stream)))

    def positions(self, seed, config):
        """Displaced positions [n_atoms, 3] of configuration ``config``."""
        out = np.zeros((self.n_atoms, 3))
        check(lib().totem_specimen_positions(self._h, ctypes.c_ulonglong(seed), int(config), _dp(out)))
        return out

    def close(self):
        if self._h:
            check(lib().totem_specimen_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
