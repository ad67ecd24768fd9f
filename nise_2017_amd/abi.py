"""ctypes view of libnise_hip.so (include/nise_abi.h).

The library is the product: this module only loads it and declares the
argument types.  Loading fails loudly when the shared object has not been
built (python -m nise_2017_amd.build).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libnise_hip.so")

NISE_OK, NISE_ERR_ARG, NISE_ERR_HIP, NISE_ERR_NOMEM, NISE_ERR_RANGE, NISE_ERR_NUMERIC = range(6)
NISE_HOST, NISE_DEVICE = 0, 1

_S256 = C.c_char * 256


class TNon(C.Structure):
    """Byte-identical mirror of the reference's t_non (src/types.h:37-90)."""

    _fields_ = [
        ("tmax1", C.c_int), ("tmax2", C.c_int), ("tmax3", C.c_int),
        ("tmin1", C.c_int), ("tmin2", C.c_int), ("tmin3", C.c_int),
        ("dt1", C.c_int), ("dt2", C.c_int), ("dt3", C.c_int),
        ("d1", C.c_int), ("d2", C.c_int), ("d3", C.c_int),
        ("min1", C.c_float), ("max1", C.c_float),
        ("min2", C.c_float), ("max2", C.c_float),
        ("min3", C.c_float), ("max3", C.c_float),
        ("k", C.c_int * 4),
        ("length", C.c_int), ("sample", C.c_int), ("fft", C.c_int),
        ("begin", C.c_int), ("end", C.c_int), ("is_", C.c_int), ("ts", C.c_int),
        ("interpol", C.c_int), ("propagation", C.c_int),
        ("lifetime", C.c_float), ("homogen", C.c_float), ("inhomogen", C.c_float),
        ("deltat", C.c_float),
        ("buffer", C.c_int), ("singles", C.c_int), ("doubles", C.c_int),
        ("energyFName", _S256), ("dipoleFName", _S256), ("alphaFName", _S256),
        ("positionFName", _S256), ("anharFName", _S256), ("overdipFName", _S256),
        ("singleShiftFName", _S256), ("technique", _S256), ("basis", _S256),
        ("pdbFName", _S256), ("hamiltonian", _S256), ("couplingFName", _S256),
        ("pbcFName", _S256),
        ("tmax", C.c_int), ("cluster", C.c_int),
        ("shifte", C.c_float), ("shiftf", C.c_float),
        ("labPol", C.c_int), ("dephasing", C.c_float),
        ("statstart", C.c_float), ("statend", C.c_float), ("statstep", C.c_float),
        ("statsteps", C.c_int), ("thres", C.c_float), ("couplingcut", C.c_float),
        ("temperature", C.c_float), ("anharmonicity", C.c_float),
        ("Npsites", C.c_int), ("psites", C.POINTER(C.c_int)),
        ("window", C.c_int),
        ("SingleShiftSites", C.c_int), ("SingleShiftSite", C.POINTER(C.c_int)),
        ("SingleShift", C.POINTER(C.c_float)),
        ("printLevel", C.c_int),
    ]


def make_non(**kw):
    """t_non with the reference's readInput defaults (src/readinput.c:25-44)
    plus the given fields; strings are encoded to char[256]."""
    n = TNon()
    n.interpol, n.begin, n.end, n.is_, n.ts = 1, 0, 0, 0, 5
    n.anharmonicity, n.couplingcut, n.temperature, n.cluster = 0.0, 0.0, 300.0, -1
    n.fft, n.printLevel, n.tmax1, n.tmax2, n.tmax3 = 0, 1, 256, 0, 256
    n.homogen, n.inhomogen, n.window = 0.0, 0.0, 0
    n.basis = b"Local"
    n.hamiltonian = b"Full"
    n.dt1 = n.dt2 = n.dt3 = 1
    keep = []
    for k, v in kw.items():
        if isinstance(v, str):
            v = v.encode()
        if k == "single_shift":  # list of (site, a, b)
            sites = (C.c_int * len(v))(*[int(s) for s, _, _ in v])
            ab = (C.c_float * (2 * len(v)))(*[float(x) for _, a, b in v for x in (a, b)])
            n.SingleShiftSites = len(v)
            n.SingleShiftSite = C.cast(sites, C.POINTER(C.c_int))
            n.SingleShift = C.cast(ab, C.POINTER(C.c_float))
            keep += [sites, ab]
            continue
        setattr(n, k, v)
    if "tmax" not in kw:
        n.tmax = n.tmax1
    n._keep = keep  # keep ctypes arrays alive with the struct
    return n


_lib = None

_fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_vp = C.c_void_p


def lib():
    """Load libnise_hip.so; raise if it is missing (no fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with python -m nise_2017_amd.build")
    # torch bundles its own libamdhip64.so.7 under the same soname as ROCm's.
    # Whichever is loaded first serves the whole process, and torch's own
    # device code needs its copy: load torch (when installed) first.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    P = C.POINTER(TNon)
    L.nise_last_error.restype = C.c_char_p
    L.nise_version.restype = C.c_char_p
    L.nise_absorption_batched.argtypes = [P, _vp, C.c_int64, _vp, C.c_int64, C.c_int64, _vp,
                                          C.c_int, C.c_int, _fp, _fp]
    L.nise_absorption_batched_noalloc.argtypes = [P, _vp, C.c_int64, _vp, C.c_int64, C.c_int64, _vp, _fp, _fp]
    L.nise_abs_plan_create.argtypes = [C.POINTER(C.c_void_p), P, _vp, C.c_int64, _vp, C.c_int64,
                                       C.c_int64, _vp, C.c_int, C.c_int]
    L.nise_abs_plan_create_onfly.argtypes = [C.POINTER(C.c_void_p), P, _vp, C.c_int64, _vp, C.c_int64,
                                             C.c_int64, _vp, C.c_int64, _vp, C.c_int, C.c_int]
    L.nise_abs_plan_execute.argtypes = [_vp, _vp, _vp, C.c_int]
    L.nise_abs_plan_set_shape.argtypes = [_vp, C.c_int, C.c_int, C.c_int]
    L.nise_abs_plan_set_rows.argtypes = [_vp, C.c_int]
    L.nise_abs_plan_autotune.argtypes = [_vp, _vp, _vp, C.c_int]
    L.nise_abs_plan_info.argtypes = [_vp, C.POINTER(C.c_int64), C.c_int]
    L.nise_abs_plan_last_timing.argtypes = [_vp, C.POINTER(C.c_float), C.c_int]
    L.nise_abs_plan_sync.argtypes = [_vp]
    L.nise_abs_plan_destroy.argtypes = [_vp]
    L.nise_abs_plan_destroy.restype = None
    L.nise_abs_plan_set_params.argtypes = [_vp, _vp, C.c_float]
    L.nise_abs_plan_enable_adjoint.argtypes = [_vp, C.c_int]
    L.nise_abs_plan_adjoint.argtypes = [_vp, _vp, _vp, _vp, C.POINTER(C.c_double)]
    L.nise_absorption_multi.argtypes = [_vp, C.c_int, C.c_int, _fp, _fp, _vp]
    L.nise_2dfft.argtypes = [P, C.c_int]
    L.nise_dft2d.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp]
    L.nise_2dir_batched.argtypes = [P, _vp, C.c_int64, _vp, C.c_int64, C.c_int64, _vp, _vp, _vp,
                                    C.c_int, _fp]
    L.nise_absorption_stream.argtypes = [P, C.c_int, _fp, _fp]
    L.nise_absorption_multi_tune.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int), C.c_void_p,
                                             C.c_void_p, C.c_int]
    L.nise_2dirraman_batched.argtypes = [P, _vp, C.c_int64, _vp, C.c_int64, C.c_int64, _vp, _vp, C.c_int,
                                         np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")]
    for name in ("propagate_vec_coupling_S",):
        getattr(L, name).argtypes = [P, _fp, _fp, _fp, C.c_int, C.c_int]
        getattr(L, name).restype = None
    L.propagate_vector.argtypes = [P, _fp, _fp, _fp, C.c_int, C.c_int, C.c_int]
    L.propagate_vector.restype = None
    L.propagate_matrix.argtypes = [P, _fp, _fp, _fp, C.c_int, C.c_int, C.c_int]
    L.propagate_matrix.restype = None
    L.propagate_vec_DIA.argtypes = [P, _fp, _fp, _fp, C.c_int]
    L.propagate_vec_DIA.restype = None
    L.diagonalizeLPD.argtypes = [_fp, _fp, C.c_int]
    L.diagonalizeLPD.restype = None
    _lib = L
    return L


def check(rc):
    if rc != NISE_OK:
        raise RuntimeError(f"libnise_hip error {rc}: {lib().nise_last_error().decode()}")


def ptr(a):
    """Device or host address of a numpy array or torch tensor."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        if not a.is_contiguous():
            raise ValueError("tensor passed to libnise_hip must be contiguous")
        return C.c_void_p(a.data_ptr())
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array passed to libnise_hip must be C-contiguous (row-major records)")
    return a.ctypes.data_as(C.c_void_p)


class AbsPlan:
    """Batched absorption plan over a trajectory (host or device records)."""

    def __init__(self, non, H_rec, H_stride, mu_rec, mu_stride, nframes, coupling_rec=None,
                 where=NISE_HOST, kernel=0, positions=None, box=None):
        """positions / box: geometry records for HamiltonianType TransitionDipole
        or ExtendedDipole (nise_abs_plan_create_onfly)."""
        L = lib()
        self._keep = (non, H_rec, mu_rec, coupling_rec, positions, box)
        h = C.c_void_p()
        if positions is not None:
            check(L.nise_abs_plan_create_onfly(C.byref(h), C.byref(non), ptr(H_rec), H_stride, ptr(mu_rec),
                                               mu_stride, nframes, ptr(positions), positions.shape[1],
                                               ptr(box), where, kernel))
        else:
            check(L.nise_abs_plan_create(C.byref(h), C.byref(non), ptr(H_rec), H_stride, ptr(mu_rec),
                                         mu_stride, nframes, ptr(coupling_rec), where, kernel))
        self.h = h
        self.tmax = non.tmax1
        self.N = non.singles

    def set_shape(self, vp, spw, block=256):
        check(lib().nise_abs_plan_set_shape(self.h, vp, spw, block))

    def execute(self, flags=0, fetch=True):
        re = np.zeros(self.tmax, np.float32)
        im = np.zeros(self.tmax, np.float32)
        if fetch:
            check(lib().nise_abs_plan_execute(self.h, ptr(re), ptr(im), flags))
            return re, im
        check(lib().nise_abs_plan_execute(self.h, None, None, flags))
        return None

    def sync(self):
        check(lib().nise_abs_plan_sync(self.h))

    def autotune(self):
        """Time the candidate configurations and keep the fastest; returns
        [(code, conveyor ms), ...] (code: rows R, or -(10000 vp + 1000 spw +
        block))."""
        ms = (C.c_float * 16)()
        codes = (C.c_int * 16)()
        k = lib().nise_abs_plan_autotune(self.h, ms, codes, 16)
        if k < 0:
            check(-k)
        return [(codes[i], ms[i]) for i in range(k)]

    def set_rows(self, rows=0):
        """Row conveyor with the given row count (0: from the structure)."""
        check(lib().nise_abs_plan_set_rows(self.h, int(rows)))

    def info(self):
        a = (C.c_int64 * 12)()
        lib().nise_abs_plan_info(self.h, a, 12)
        keys = ("N", "K", "nlayers", "nwg", "vp", "spw", "f_lo", "nframes", "block", "nsteps",
                "rows", "rotations")
        return dict(zip(keys, list(a)))

    def timing(self):
        a = (C.c_float * 4)()
        lib().nise_abs_plan_last_timing(self.h, a, 4)
        return {"tables_ms": a[0], "conveyor_ms": a[1], "reduce_ms": a[2], "adjoint_ms": a[3]}

    def set_params(self, delta=None, cscale=1.0):
        """Per-site energy offsets (N floats or None) and coupling scale."""
        d = None
        if delta is not None:
            d = np.ascontiguousarray(delta, np.float32)
            if d.shape != (self.N,):
                raise ValueError(f"delta must have shape ({self.N},)")
        check(lib().nise_abs_plan_set_params(self.h, ptr(d), float(cscale)))

    def enable_adjoint(self, every):
        check(lib().nise_abs_plan_enable_adjoint(self.h, int(every)))

    def adjoint(self, g_re, g_im):
        """(dL/d delta (N,), dL/d cscale) of the last execute for
        g = dL/dRe S + i dL/dIm S."""
        gr = np.ascontiguousarray(g_re, np.float32)
        gi = np.ascontiguousarray(g_im, np.float32)
        if gr.shape != (self.tmax,) or gi.shape != (self.tmax,):
            raise ValueError(f"gradient inputs must have shape ({self.tmax},)")
        gd = np.zeros(self.N, np.float64)
        gc = C.c_double()
        check(lib().nise_abs_plan_adjoint(self.h, ptr(gr), ptr(gi), ptr(gd), C.byref(gc)))
        return gd, gc.value

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.nise_abs_plan_destroy(self.h)
        self.h = None

    def __del__(self):
        self.close()


class NiseSystem(C.Structure):
    """Mirror of nise_system (include/nise_abi.h)."""
    _fields_ = [("non", C.POINTER(TNon)), ("H_rec", C.c_void_p), ("H_stride", C.c_int64),
                ("mu_rec", C.c_void_p), ("mu_stride", C.c_int64), ("nframes", C.c_int64),
                ("coupling_rec", C.c_void_p)]


def absorption_multi(systems, where=None):
    """nise_absorption_multi over [(non, H, mu, coupling_rec or None), ...]:
    returns ([(re_S, im_S) per system], done_ms array)."""
    if where is None:
        H0 = systems[0][1]
        where = NISE_DEVICE if hasattr(H0, "data_ptr") and H0.is_cuda else NISE_HOST
    arr = (NiseSystem * len(systems))()
    keep = []
    for k, (non, H, mu, Cr) in enumerate(systems):
        for a in (H, mu, Cr):
            if a is not None and not hasattr(a, "data_ptr") and a.dtype != np.float32:
                raise ValueError("records must be float32")
        arr[k] = NiseSystem(C.pointer(non), ptr(H), H.shape[1], ptr(mu), mu.shape[1], H.shape[0],
                            ptr(Cr))
        keep.append(non)
    T = [non.tmax1 for non, _, _, _ in systems]
    re = np.zeros(sum(T), np.float32)
    im = np.zeros(sum(T), np.float32)
    done = np.zeros(len(systems), np.float64)
    check(lib().nise_absorption_multi(C.cast(arr, C.c_void_p), len(systems), where, re, im,
                                      done.ctypes.data_as(C.c_void_p)))
    out, o = [], 0
    for t in T:
        out.append((re[o:o + t].copy(), im[o:o + t].copy()))
        o += t
    return out, done


def absorption_multi_tune(systems, where=None):
    """nise_absorption_multi_tune over the same list as absorption_multi:
    returns (row count kept, [(row count, ms per call), ...])."""
    if where is None:
        H0 = systems[0][1]
        where = NISE_DEVICE if hasattr(H0, "data_ptr") and H0.is_cuda else NISE_HOST
    arr = (NiseSystem * len(systems))()
    keep = []
    for k, (non, H, mu, Cr) in enumerate(systems):
        arr[k] = NiseSystem(C.pointer(non), ptr(H), H.shape[1], ptr(mu), mu.shape[1], H.shape[0], ptr(Cr))
        keep.append(non)
    best = C.c_int()
    ms = (C.c_float * 16)()
    codes = (C.c_int * 16)()
    k = lib().nise_absorption_multi_tune(C.cast(arr, C.c_void_p), len(systems), where, C.byref(best), ms, codes, 16)
    if k < 0:
        check(-k)
    return best.value, [(codes[i], ms[i]) for i in range(min(k, 16))]


def twodir_raman_batched(non, H, mu, alpha, nframes=None, coupling_rec=None, where=None):
    """nise_2dirraman_batched: the 12 2DIR-Raman planes (12, tmax3, tmax1)
    float32 summed over samples and polarisation items (un-normalised);
    alpha holds the Raman tensor records (frames, 1 + 6N)."""
    if where is None:
        where = NISE_DEVICE if hasattr(H, "data_ptr") and H.is_cuda else NISE_HOST
    if where == NISE_HOST:
        for a in (H, mu, alpha, coupling_rec):
            if a is not None and a.dtype != np.float32:
                raise ValueError("records must be float32")
    R = np.zeros((12, non.tmax3, non.tmax1), np.float32)
    nfr = H.shape[0] if nframes is None else nframes
    check(lib().nise_2dirraman_batched(C.byref(non), ptr(H), H.shape[1], ptr(mu), mu.shape[1], nfr,
                                       ptr(coupling_rec), ptr(alpha), where, R))
    return R


def twodir_batched(non, H, mu, nframes=None, coupling_rec=None, anh_rec=None, over_rec=None,
                   where=None):
    """nise_2dir_batched: the 12 response planes (12, tmax3, tmax1) float32
    summed over samples and polarisation items (un-normalised).  Records
    are numpy arrays (host) or contiguous device tensors (where inferred)."""
    if where is None:
        where = NISE_DEVICE if hasattr(H, "data_ptr") and H.is_cuda else NISE_HOST
    if where == NISE_HOST:
        for a in (H, mu, coupling_rec, anh_rec, over_rec):
            if a is not None and a.dtype != np.float32:
                raise ValueError("records must be float32")
    R = np.zeros((12, non.tmax3, non.tmax1), np.float32)
    nfr = H.shape[0] if nframes is None else nframes
    check(lib().nise_2dir_batched(C.byref(non), ptr(H), H.shape[1], ptr(mu), mu.shape[1], nfr,
                                  ptr(coupling_rec), ptr(anh_rec), ptr(over_rec), where, R))
    return R
