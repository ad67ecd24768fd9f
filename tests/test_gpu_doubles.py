"""GPU parity: two-exciton (doubles) Trotter propagation and the 2DIR
propagation control against the oracle (reference order, FP32)."""
import ctypes as C

import numpy as np
import pytest

from nise_2017_amd import abi, synth
from nise_2017_amd.abi import make_non

pytestmark = pytest.mark.gpu


def frame(rng, N, density=1.0):
    nn = N * (N + 1) // 2
    H = rng.normal(0, 30, nn).astype(np.float32)
    if density < 1:
        H[rng.random(nn) > density] = 0
    for a in range(N):
        H[synth.sindex(a, a, N)] = rng.normal(0, 50)
    return H


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def L():
    L = abi.lib()
    P = C.POINTER(abi.TNon)
    fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    L.propagate_vec_coupling_S_doubles.argtypes = [P, fp, fp, fp, C.c_int, fp]
    L.propagate_vec_coupling_S_doubles.restype = None
    pp = C.POINTER(C.POINTER(C.c_float))
    L.propagate_double_2DIR_control.argtypes = [P, fp, pp, pp, fp, fp, pp, pp, fp, fp, C.c_int,
                                                C.c_int, C.c_int, fp]
    L.propagate_double_2DIR_control.restype = None
    return L


@pytest.mark.parametrize("N,m,anh,cut,density", [
    (1, 1, 16.0, 0.0, 1.0), (2, 1, 16.0, 0.0, 1.0), (5, 2, 0.0, 0.0, 1.0), (8, 1, 20.0, 5.0, 1.0),
    (17, 1, 16.0, 0.0, 0.4), (32, 1, 16.0, 0.0, 0.2), (40, 3, 0.0, 1.0, 0.1), (70, 1, 16.0, 0.0, 0.02)])
def test_doubles_matches_reference_order(orc, L, N, m, anh, cut, density):
    rng = np.random.default_rng(100 + N)
    H = frame(rng, N, density)
    N2 = N * (N + 1) // 2
    Anh = rng.normal(16, 2, N).astype(np.float32)
    fr = rng.normal(size=N2).astype(np.float32)
    fi = rng.normal(size=N2).astype(np.float32)
    non = make_non(singles=N, deltat=2.0, couplingcut=cut, anharmonicity=anh)
    g_r, g_i, o_r, o_i = fr.copy(), fi.copy(), fr.copy(), fi.copy()
    L.propagate_vec_coupling_S_doubles(C.byref(non), H, g_r, g_i, m, Anh)
    orc.oracle().or_coupling_S_doubles_f32(N, H, o_r, o_i, m, Anh, anh, 2.0, cut)
    # float libm (reference) vs correctly rounded trig (here): a few ulp
    assert rel(g_r + 1j * g_i, o_r + 1j * o_i) < 1e-6
    for _ in range(4):
        L.propagate_vec_coupling_S_doubles(C.byref(non), H, g_r, g_i, m, Anh)
        orc.oracle().or_coupling_S_doubles_f32(N, H, o_r, o_i, m, Anh, anh, 2.0, cut)
    assert rel(g_r + 1j * g_i, o_r + 1j * o_i) < 3e-6
    d_r, d_i = fr.astype(np.float64), fi.astype(np.float64)
    for _ in range(5):
        orc.oracle().or_coupling_S_doubles_f64(N, H, d_r, d_i, m, Anh, anh, 2.0, cut)
    assert rel(g_r + 1j * g_i, d_r + 1j * d_i) < 3e-5


def test_2dir_control(orc, L):
    rng = np.random.default_rng(7)
    N, T1 = 9, 5
    N2 = N * (N + 1) // 2
    H = frame(rng, N)
    Anh = np.zeros(N, np.float32)
    non = make_non(singles=N, deltat=2.0, anharmonicity=16.0, tmax1=T1, ts=1, propagation=1)
    ft = rng.normal(size=(T1, 2, N2)).astype(np.float32)
    f0 = rng.normal(size=(2, N2)).astype(np.float32)
    rt = rng.normal(size=(T1, 2, N)).astype(np.float32)
    r0 = rng.normal(size=(2, N)).astype(np.float32)
    g_ft, g_f0, g_rt, g_r0 = ft.copy(), f0.copy(), rt.copy(), r0.copy()

    def ptrs(arr, k):
        return (C.POINTER(C.c_float) * T1)(*[arr[t, k].ctypes.data_as(C.POINTER(C.c_float))
                                             for t in range(T1)])
    L.propagate_double_2DIR_control(C.byref(non), H, ptrs(g_ft, 0), ptrs(g_ft, 1), g_f0[0], g_f0[1],
                                    ptrs(g_rt, 0), ptrs(g_rt, 1), g_r0[0], g_r0[1], 0, 0, 0, Anh)
    O = orc.oracle()
    o = f0.copy()
    O.or_coupling_S_doubles_f32(N, H, o[0], o[1], 1, Anh, 16.0, 2.0, 0.0)
    assert rel(g_f0, o) < 1e-6
    for t in range(T1):
        o = ft[t].copy()
        O.or_coupling_S_doubles_f32(N, H, o[0], o[1], 1, Anh, 16.0, 2.0, 0.0)
        assert rel(g_ft[t], o) < 1e-6
        o = rt[t].copy()
        O.or_coupling_S_f32(N, H, o[0], o[1], 1, -1, 2.0, 0.0)
        assert rel(g_rt[t], o) < 1e-6
    o = r0.copy()
    O.or_coupling_S_f32(N, H, o[0], o[1], 1, -1, 2.0, 0.0)
    assert rel(g_r0, o) < 1e-6
