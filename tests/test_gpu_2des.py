"""GPU parity of the two-level-site 2D response (calc_2DES, 2DUVvis) through
the C ABI: the batched response (nise_2dir_batched with the UVvis
techniques), the two-level doubles step (propagate_vec_coupling_S_doubles_ES),
its control routine (propagate_double_2DES_control) and the calc_2DES file
drop-in, against the oracle (restating src/calc_2DES.c:22-586 and
src/propagate.c:1051-1181, bit-identical to the reference build on the
primitives) and the closed form of uncoupled two-level sites.

Tolerances as for 2DIR: 1e-4 relative L2 over the 12 planes (FP32
propagation in the reference order, sums over samples and items in FP64
here and FP32 in the oracle); the doubles step within a few ulp (float libm
in the reference, correctly rounded trigonometry here).
"""
import ctypes as C
import os

import numpy as np
import pytest

from nise_2017_amd import abi, synth
from nise_2017_amd.abi import make_non

import twodir_closed as CF

pytestmark = pytest.mark.gpu

TOL = 1e-4


def rel(a, b):
    a = np.asarray(a)
    a = a.astype(np.result_type(a.dtype, np.float64))  # keeps complex inputs complex
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def say(msg):
    print(f"[2des] {msg}", flush=True)


def mk(N, T1, T3, sr, begin, end, H, **kw):
    lo, hi = synth.shifted_center(H, N)
    base = dict(singles=N, tmax1=T1, tmax2=0, tmax3=T3, sample=sr, begin=begin, end=end, ts=1,
                propagation=1, deltat=2.0, min1=lo, max1=hi, couplingcut=0.0, technique="2DUVvis",
                anharmonicity=16.0, lifetime=0.0)
    base.update(kw)
    return make_non(**base)


@pytest.mark.parametrize("N,T1,T3,tech", [(6, 8, 8, "2DUVvis"), (5, 6, 9, "GBUVvis"), (5, 6, 9, "SEUVvis"),
                                          (5, 6, 9, "EAUVvis"), (5, 6, 9, "noEAUVvis"),
                                          (32, 13, 5, "2DUVvis")])
def test_batched_vs_oracle(orc, N, T1, T3, tech):
    say(f"batched {tech} N={N} T1={T1} T3={T3}")
    H, mu, anh, over = synth.twodir_system(N=N, nframes=80, seed=71 + N)
    non = mk(N, T1, T3, 7, 1, 4, H, technique=tech, lifetime=400.0)
    ref = orc.twodir(non, H, mu)
    non = mk(N, T1, T3, 7, 1, 4, H, technique=tech, lifetime=400.0)
    R = abi.twodir_batched(non, H, mu)
    assert rel(R, ref) < TOL


def test_wide_system_vs_oracle(orc):
    """N = 160: the unstaged EA path with the two-level contraction."""
    say("batched 2DUVvis N=160")
    N, T1, T3 = 160, 3, 3
    H, mu, anh, over = synth.twodir_system(N=N, nframes=30, seed=72)
    non = mk(N, T1, T3, 9, 0, 2, H)
    ref = orc.twodir(non, H, mu, acc64=True)
    non = mk(N, T1, T3, 9, 0, 2, H)
    assert rel(abi.twodir_batched(non, H, mu), ref) < TOL


def test_uncoupled_two_level_closed_form():
    say("uncoupled two-level sites, closed form")
    N, T1, T3 = 3, 7, 5
    H, mu, anh, over = synth.twodir_system(N=N, nframes=60, coupled=False, seed=73)
    non = mk(N, T1, T3, 4, 2, 6, H, lifetime=250.0)
    R = abi.twodir_batched(non, H, mu)
    assert rel(R, CF.two_level(non, H, mu, es=True)) < TOL


def test_anharmonicity_is_ignored():
    """The two-level forms read no anharmonicity (records or fixed)."""
    say("anharmonicity ignored")
    N, T1, T3 = 4, 5, 5
    H, mu, anh, over = synth.twodir_system(N=N, nframes=60, seed=74)
    a = abi.twodir_batched(mk(N, T1, T3, 5, 0, 4, H, anharmonicity=16.0), H, mu)
    b = abi.twodir_batched(mk(N, T1, T3, 5, 0, 4, H, anharmonicity=0.0), H, mu)
    assert np.array_equal(a, b)


def frame(rng, N, density=1.0):
    nn = N * (N + 1) // 2
    H = rng.normal(0, 30, nn).astype(np.float32)
    if density < 1:
        H[rng.random(nn) > density] = 0
    for a in range(N):
        H[synth.sindex(a, a, N)] = rng.normal(0, 50)
    return H


@pytest.fixture(scope="module")
def L():
    L = abi.lib()
    P = C.POINTER(abi.TNon)
    fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    L.propagate_vec_coupling_S_doubles_ES.argtypes = [P, fp, fp, fp, C.c_int]
    L.propagate_vec_coupling_S_doubles_ES.restype = None
    pp = C.POINTER(C.POINTER(C.c_float))
    L.propagate_double_2DES_control.argtypes = [P, fp, pp, pp, fp, fp, pp, pp, fp, fp, C.c_int, C.c_int,
                                                C.c_int]
    L.propagate_double_2DES_control.restype = None
    return L


@pytest.mark.parametrize("N,m,cut,density", [(1, 1, 0.0, 1.0), (2, 1, 0.0, 1.0), (5, 2, 0.0, 1.0),
                                             (8, 1, 5.0, 1.0), (32, 1, 0.0, 0.2), (70, 3, 1.0, 0.05)])
def test_doubles_es_matches_reference_order(orc, L, N, m, cut, density):
    say(f"doubles ES N={N} m={m}")
    rng = np.random.default_rng(200 + N)
    H = frame(rng, N, density)
    N2 = N * (N + 1) // 2
    fr = rng.normal(size=N2).astype(np.float32)
    fi = rng.normal(size=N2).astype(np.float32)
    non = make_non(singles=N, deltat=2.0, couplingcut=cut, anharmonicity=16.0)
    g_r, g_i, o_r, o_i = fr.copy(), fi.copy(), fr.copy(), fi.copy()
    for _ in range(5):
        L.propagate_vec_coupling_S_doubles_ES(C.byref(non), H, g_r, g_i, m)
        orc.oracle().or_coupling_S_doubles_ES_f32(N, H, o_r, o_i, m, 2.0, cut)
    assert rel(g_r + 1j * g_i, o_r + 1j * o_i) < 3e-6
    # the bosonic step differs whenever a coupling is kept
    if N > 1:
        b_r, b_i = fr.copy(), fi.copy()
        orc.oracle().or_coupling_S_doubles_f32(N, H, b_r, b_i, m, np.zeros(N, np.float32), 16.0, 2.0, cut)
        g_r, g_i = fr.copy(), fi.copy()
        L.propagate_vec_coupling_S_doubles_ES(C.byref(non), H, g_r, g_i, m)
        assert rel(g_r + 1j * g_i, b_r + 1j * b_i) > 1e-4


def test_2des_control(orc, L):
    say("propagate_double_2DES_control")
    rng = np.random.default_rng(9)
    N, T1 = 9, 5
    N2 = N * (N + 1) // 2
    H = frame(rng, N)
    non = make_non(singles=N, deltat=2.0, anharmonicity=16.0, tmax1=T1, ts=1, propagation=1)
    ft = rng.normal(size=(T1, 2, N2)).astype(np.float32)
    f0 = rng.normal(size=(2, N2)).astype(np.float32)
    rt = rng.normal(size=(T1, 2, N)).astype(np.float32)
    r0 = rng.normal(size=(2, N)).astype(np.float32)
    g_ft, g_f0, g_rt, g_r0 = ft.copy(), f0.copy(), rt.copy(), r0.copy()
    PF = C.POINTER(C.c_float)

    def rows(a, k):
        return (PF * T1)(*[a[t, k].ctypes.data_as(PF) for t in range(T1)])

    L.propagate_double_2DES_control(C.byref(non), H, rows(g_ft, 0), rows(g_ft, 1), g_f0[0], g_f0[1],
                                    rows(g_rt, 0), rows(g_rt, 1), g_r0[0], g_r0[1], 0, 0, 0)
    o = orc.oracle()
    o_ft, o_f0, o_rt, o_r0 = ft.copy(), f0.copy(), rt.copy(), r0.copy()
    o.or_coupling_S_doubles_ES_f32(N, H, o_f0[0], o_f0[1], 1, 2.0, 0.0)
    for t in range(T1):
        o.or_coupling_S_doubles_ES_f32(N, H, o_ft[t, 0], o_ft[t, 1], 1, 2.0, 0.0)
        o.or_coupling_S_f32(N, H, o_rt[t, 0], o_rt[t, 1], 1, -1, 2.0, 0.0)
    o.or_coupling_S_f32(N, H, o_r0[0], o_r0[1], 1, -1, 2.0, 0.0)
    for g, e in ((g_ft, o_ft), (g_f0, o_f0), (g_rt, o_rt), (g_r0, o_r0)):
        assert rel(g, e) < 1e-6


def test_calc_2des_files(orc, tmp_path, monkeypatch):
    say("calc_2DES files")
    Lb = abi.lib()
    Lb.calc_2DES.argtypes = [C.POINTER(abi.TNon), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    Lb.calc_2DES.restype = None
    N, T1, T3, sr = 5, 6, 6, 5
    H, mu, anh, over = synth.twodir_system(N=N, nframes=70, seed=75)
    d = str(tmp_path)
    H.tofile(os.path.join(d, "H.bin")), mu.tofile(os.path.join(d, "mu.bin"))
    # no anharmonicity files: calc_2DES does not read them
    kw = dict(length=H.shape[0], anharmonicity=0.0, lifetime=300.0,
              energyFName=os.path.join(d, "H.bin"), dipoleFName=os.path.join(d, "mu.bin"))
    non = mk(N, T1, T3, sr, 2, 0, H, **kw)
    total = (H.shape[0] - T1 - T3 - 1) // sr + 1
    ref = orc.twodir(mk(N, T1, T3, sr, 2, total, H, **kw), H, mu)
    monkeypatch.chdir(d)
    Lb.calc_2DES(C.byref(non), 0, 1, 0, 1, 0, 0)
    samples = total - 2
    names = ["RparI.dat", "RparII.dat", "RperI.dat", "RperII.dat", "RcroI.dat", "RcroII.dat"]
    for p in range(3):
        for k in range(2):
            re = ref[4 * p + 2 * k].astype(np.float32) / np.float32(samples)
            im = ref[4 * p + 2 * k + 1].astype(np.float32) / np.float32(samples)
            re[0] /= 2
            im[0] /= 2
            e = np.array([[t1 * 2.0, 0.0, t3 * 2.0, re[t3, t1], im[t3, t1]]
                          for t1 in range(T1) for t3 in range(T3)])
            got = np.loadtxt(os.path.join(d, names[2 * p + k]), ndmin=2)
            assert got.shape == e.shape and np.allclose(got[:, :3], e[:, :3])
            assert rel(got[:, 3:], e[:, 3:]) < 1e-4, names[2 * p + k]
