"""ORACLE -- TEST INFRASTRUCTURE ONLY.

ctypes access to the CPU restatement (oracle/_build/libnise_oracle.so) and,
where it was built, to the reference-source driver (oracle/_ref/libnise_ref.so).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "libnise_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libnise_ref.so")
REF_IEEE_SO = os.path.join(HERE, "_ref", "libnise_ref_ieee.so")  # ACCURATE_MATHS flags

_f32 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64 = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32 = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_vp = C.c_void_p


def build(ref=None):
    """Compile the oracle (and the reference driver when the reference
    sources are present)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if ref is None:
        ref = os.path.isdir("/root/reference/src")
    if ref:
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_o = None
_r = None
_ri = None


def oracle():
    global _o
    if _o is not None:
        return _o
    from nise_2017_amd.abi import TNon
    P = C.POINTER(TNon)
    L = C.CDLL(ORACLE_SO)
    L.or_sindex.argtypes = [C.c_int] * 3
    L.or_coupling_S_f32.argtypes = [C.c_int, _f32, _f32, _f32, C.c_int, C.c_int, C.c_float, C.c_float]
    L.or_coupling_S_f64.argtypes = [C.c_int, _f32, _f64, _f64, C.c_int, C.c_int, C.c_float, C.c_float]
    L.or_coupling_S_inv_f64.argtypes = L.or_coupling_S_f64.argtypes
    L.or_coupling_S_doubles_f32.argtypes = [C.c_int, _f32, _f32, _f32, C.c_int, _f32, C.c_float,
                                            C.c_float, C.c_float]
    L.or_coupling_S_doubles_f64.argtypes = [C.c_int, _f32, _f64, _f64, C.c_int, _f32, C.c_float,
                                            C.c_float, C.c_float]
    L.or_jacobi_f64.argtypes = [C.c_int, _f64, _f64, _f64]
    L.or_DIA_f64.argtypes = [C.c_int, _f32, _f64, _f64, C.c_int, C.c_float]
    L.or_DIA_S_f64.argtypes = [C.c_int, _f32, _f64, _f64, C.c_int, C.c_float, C.c_float]
    L.or_DIA_S_f64.restype = C.c_int
    L.or_set_geometry.argtypes = [_vp, C.c_longlong, _vp, C.c_longlong, _vp]
    L.or_dipole_double.argtypes = [C.c_int, _f32, _f32, _f32, _f32, _f32, _f32, C.c_float]
    L.or_dipole_double_last.argtypes = L.or_dipole_double.argtypes
    L.or_coupling_S_doubles_ES_f32.argtypes = [C.c_int, _f32, _f32, _f32, C.c_int, C.c_float, C.c_float]
    L.or_dipole_double_ES.argtypes = [C.c_int, _f32, _f32, _f32, _f32, _f32]
    L.or_dipole_double_last_ES.argtypes = L.or_dipole_double_ES.argtypes
    L.or_polarweight.argtypes = [C.c_int, C.c_int]
    L.or_polarweight.restype = C.c_float
    L.or_polar.argtypes = [_i32, C.c_int]
    L.or_read_H.argtypes = [P, _vp, C.c_longlong, _vp, C.c_longlong, _f32, C.c_int]
    L.or_absorption.argtypes = [P, _vp, C.c_longlong, _vp, C.c_longlong, _vp, _vp, _vp, _vp, _vp,
                                C.c_int]
    L.or_1dfft.argtypes = [P, _f32, _f32, C.c_int, _f64, _f64, _f64, C.POINTER(C.c_int)]
    L.or_absorption_param_f64.argtypes = [P, _vp, C.c_longlong, _vp, C.c_longlong, _vp, _vp,
                                          C.c_double, _f64, _f64, C.c_int]
    L.or_time_evolution_mat.argtypes = [C.c_int, _f32, C.c_float, C.c_int, C.c_float, _f32, _f32, _i32, _i32]
    L.or_double_sparce_f32.argtypes = [C.c_int, _f32, _f32, _i32, _i32, _f32, _f32, C.c_int, C.c_int, _vp,
                                       C.c_float, C.c_float, C.c_int]
    L.or_2dir_item.argtypes = [P, _vp, C.c_longlong, _vp, C.c_longlong, _vp, _vp, _vp, C.c_int,
                               C.c_int, _f32]
    L.or_dipole_double_ground.argtypes = [C.c_int, _f32, _f32, _f32, _f32, C.c_float]
    L.or_raman_pol.argtypes = [C.c_int, C.c_int]
    L.or_2dirraman_item.argtypes = [P, _vp, C.c_longlong, _vp, C.c_longlong, _vp, _vp, C.c_int, C.c_int,
                                    _f32]
    _o = L
    return L


def have_ref():
    return os.path.exists(REF_SO)


def ref(ieee=False):
    """The reference build (default flags, -ffast-math), or with ieee=True
    the same sources built with the reference's ACCURATE_MATHS flags."""
    global _r, _ri
    if ieee:
        if _ri is None:
            _ri = _bind_ref(C.CDLL(REF_IEEE_SO))
        return _ri
    if _r is None:
        _r = _bind_ref(C.CDLL(REF_SO))
    return _r


def _bind_ref(L):
    L.ref_coupling_S.argtypes = [C.c_int, _f32, _f32, _f32, C.c_int, C.c_int, C.c_float, C.c_float]
    L.ref_coupling_S_doubles.argtypes = [C.c_int, _f32, _f32, _f32, C.c_int, _f32, C.c_float,
                                         C.c_float, C.c_float]
    L.ref_dipole_double.argtypes = [C.c_int, _f32, _f32, _f32, _f32, _f32, _f32, C.c_float]
    L.ref_dipole_double_last.argtypes = L.ref_dipole_double.argtypes
    L.ref_coupling_S_doubles_ES.argtypes = [C.c_int, _f32, _f32, _f32, C.c_int, C.c_float, C.c_float]
    L.ref_dipole_double_ES.argtypes = [C.c_int, _f32, _f32, _f32, _f32, _f32]
    L.ref_dipole_double_last_ES.argtypes = L.ref_dipole_double_ES.argtypes
    L.ref_polarweight.argtypes = [C.c_int, C.c_int]
    L.ref_polarweight.restype = C.c_float
    L.ref_polar.argtypes = [_i32, C.c_int]
    L.ref_sindex.argtypes = [C.c_int] * 3
    L.ref_tnon_layout.argtypes = [np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS"), C.c_int]
    cs = C.c_char_p
    L.ref_read_frame.argtypes = [C.c_int, cs, cs, cs, cs, cs, C.c_float, C.c_float, C.c_int, _f32, _f32]
    L.ref_absorption_files.argtypes = [C.c_int, cs, cs, cs, cs, cs, C.c_float, C.c_float, C.c_float,
                                       C.c_float, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.c_int, _f32, _f32]
    L.ref_read_frame_onfly.argtypes = [C.c_int, cs, cs, cs, cs, cs, C.c_float, C.c_float, C.c_int, _f32]
    L.ref_absorption_mem.argtypes = [C.c_int, _vp, C.c_longlong, _vp, C.c_longlong, C.c_float, C.c_float,
                                     C.c_float, C.c_float, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_int, _f32, _f32]
    L.ref_absorption_files_geo.argtypes = [C.c_int, cs, cs, cs, cs, cs, cs, cs, C.c_float, C.c_float,
                                           C.c_float, C.c_float, C.c_int, C.c_int, C.c_int, C.c_int,
                                           C.c_int, C.c_int, _f32, _f32]
    if hasattr(L, "ref_dipole_double_ground"):
        L.ref_dipole_double_ground.argtypes = [C.c_int, _f32, _f32, _f32, _f32, C.c_float]
        L.ref_find_raman_pol.argtypes = [C.c_int, C.c_int]
        L.ref_read_alpha.argtypes = [C.c_int, cs, C.c_int, C.c_int, _f32]
    return L


def ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(C.c_void_p)


def absorption(non, H, mu, C_rec=None, f64=False, nthreads=1):
    """Oracle absorption: returns (re_S, im_S) float32 or float64 arrays."""
    T = non.tmax1
    H = np.ascontiguousarray(H, np.float32)
    mu = np.ascontiguousarray(mu, np.float32)
    Cr = None if C_rec is None else np.ascontiguousarray(C_rec, np.float32)
    if f64:
        re = np.zeros(T, np.float64)
        im = np.zeros(T, np.float64)
        rc = oracle().or_absorption(C.byref(non), ptr(H), H.shape[1], ptr(mu), mu.shape[1], ptr(Cr),
                                    None, None, ptr(re), ptr(im), nthreads)
    else:
        re = np.zeros(T, np.float32)
        im = np.zeros(T, np.float32)
        rc = oracle().or_absorption(C.byref(non), ptr(H), H.shape[1], ptr(mu), mu.shape[1], ptr(Cr),
                                    ptr(re), ptr(im), None, None, nthreads)
    assert rc == 0
    return re, im


def fft1d(non, re_S, im_S, samples):
    n = max(non.fft, non.tmax1)
    fr = np.zeros(n, np.float64)
    sr = np.zeros(n, np.float64)
    si = np.zeros(n, np.float64)
    cnt = C.c_int()
    oracle().or_1dfft(C.byref(non), np.ascontiguousarray(re_S, np.float32),
                      np.ascontiguousarray(im_S, np.float32), samples, fr, sr, si, C.byref(cnt))
    k = cnt.value
    return fr[:k], sr[:k], si[:k]


def twodir(non, H, mu, C_rec=None, A_rec=None, over_rec=None, acc64=False):
    """Oracle 2DIR: the 12 response planes (12, tmax3, tmax1) float32 summed
    over samples begin..end-1 and the 21 polarisation items in the
    reference's work order (src/MPI_subs.c:74-95), un-normalised.  Sets
    non.shifte = (min1 + max1) / 2 as src/calc_2DIR.c:79-81 does.

    acc64: each item's contribution is formed in FP32 as the reference does
    and the items are summed in FP64 (float64 planes), which removes the
    rounding of the reference's FP32 running sums over items."""
    H = np.ascontiguousarray(H, np.float32)
    mu = np.ascontiguousarray(mu, np.float32)
    Cr = None if C_rec is None else np.ascontiguousarray(C_rec, np.float32)
    Ar = None if A_rec is None else np.ascontiguousarray(A_rec, np.float32)
    Or = None if over_rec is None else np.ascontiguousarray(over_rec, np.float32)
    non.shifte = np.float32((np.float32(non.max1) + np.float32(non.min1)) / np.float32(2))
    R = np.zeros((12, non.tmax3, non.tmax1), np.float32)
    R64 = np.zeros(R.shape, np.float64) if acc64 else None
    end = non.end
    if end == 0:
        end = (H.shape[0] - non.tmax1 - non.tmax2 - non.tmax3 - 1) // non.sample + 1
    for s in range(non.begin, end):
        for m in range(21):
            if acc64:
                R[...] = 0
            rc = oracle().or_2dir_item(C.byref(non), ptr(H), H.shape[1], ptr(mu), mu.shape[1],
                                       ptr(Cr), ptr(Ar), ptr(Or), s, m, R)
            assert rc == 0
            if acc64:
                R64 += R
    return R64 if acc64 else R


def twodir_raman(non, H, mu, alpha, C_rec=None, acc64=False):
    """Oracle 2DIR-Raman: the 12 response planes (12, tmax3, tmax1) summed over
    samples begin..end-1 and the 21 items in the reference's work order, as
    twodir(); alpha holds the Raman tensor records (frames, 1 + 6N)."""
    H = np.ascontiguousarray(H, np.float32)
    mu = np.ascontiguousarray(mu, np.float32)
    al = np.ascontiguousarray(alpha, np.float32)
    Cr = None if C_rec is None else np.ascontiguousarray(C_rec, np.float32)
    non.shifte = np.float32((np.float32(non.max1) + np.float32(non.min1)) / np.float32(2))
    R = np.zeros((12, non.tmax3, non.tmax1), np.float32)
    R64 = np.zeros(R.shape, np.float64) if acc64 else None
    end = non.end
    if end == 0:
        end = (H.shape[0] - non.tmax1 - non.tmax2 - non.tmax3 - 1) // non.sample + 1
    for s in range(non.begin, end):
        for m in range(21):
            if acc64:
                R[...] = 0
            rc = oracle().or_2dirraman_item(C.byref(non), ptr(H), H.shape[1], ptr(mu), mu.shape[1], ptr(Cr),
                                            ptr(al), s, m, R)
            assert rc == 0
            if acc64:
                R64 += R
    return R64 if acc64 else R


def absorption_param_f64(non, H, mu, delta=None, cscale=1.0, C_rec=None, nthreads=1):
    """FP64 oracle absorption with per-site offsets delta (N) and coupling
    scale cscale: returns (re_S, im_S) float64."""
    T = non.tmax1
    H = np.ascontiguousarray(H, np.float32)
    mu = np.ascontiguousarray(mu, np.float32)
    Cr = None if C_rec is None else np.ascontiguousarray(C_rec, np.float32)
    d = None if delta is None else np.ascontiguousarray(delta, np.float64)
    re = np.zeros(T, np.float64)
    im = np.zeros(T, np.float64)
    rc = oracle().or_absorption_param_f64(C.byref(non), ptr(H), H.shape[1], ptr(mu), mu.shape[1],
                                          ptr(Cr), ptr(d), float(cscale), re, im, nthreads)
    assert rc == 0
    return re, im
