"""The reference's own absorption loop as the CPU baseline times it
(oracle/_ref: src/absorption.c:92-146 over the reference's read_He,
read_mue and propagate_vec_coupling_S): reading the trajectory through
stdio streams over memory gives the same sums, bit for bit, as reading the
files; sample ranges run on separate threads add up to the whole run (to
FP32 rounding of the split running sum); and the loop agrees with the
oracle's absorption."""
import os
import threading

import numpy as np
import pytest

from nise_2017_amd import abi, synth


@pytest.fixture(scope="module")
def ref(orc):
    if not orc.have_ref():
        pytest.skip("reference sources not built here")
    return orc.ref()


def test_memory_streams_equal_files(orc, ref, tmp_path):
    N, T, sr, ns = 9, 24, 5, 8
    rng, mus, coup = synth.lattice_static(3, seed=91)
    nfr = (ns - 1) * sr + T + 2
    H = synth.full_records(synth.lattice_diag(rng, nfr, N), coup)
    mu = synth.dipole_records(mus, nfr)
    lo, hi = synth.shifted_center(H, N)
    hf, df = str(tmp_path / "H.bin"), str(tmp_path / "mu.bin")
    H.tofile(hf)
    mu.tofile(df)
    a = (np.zeros(T, np.float32), np.zeros(T, np.float32))
    b = (np.zeros(T, np.float32), np.zeros(T, np.float32))
    ref.ref_absorption_files(N, hf.encode(), df.encode(), b"", b"", b"Full", lo, hi, 2.0, 0.0, 1, T, sr, 0, ns,
                             nfr, *a)
    ref.ref_absorption_mem(N, H.ctypes.data, H.nbytes, mu.ctypes.data, mu.nbytes, lo, hi, 2.0, 0.0, 1, T, sr, 0,
                           ns, nfr, *b)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    # sample ranges on threads
    parts = [(np.zeros(T, np.float32), np.zeros(T, np.float32)) for _ in range(3)]
    rng_ = [(0, 3), (3, 5), (5, 8)]
    ths = [threading.Thread(target=ref.ref_absorption_mem,
                            args=(N, H.ctypes.data, H.nbytes, mu.ctypes.data, mu.nbytes, lo, hi, 2.0, 0.0, 1, T, sr,
                                  s0, s1, nfr, *parts[i])) for i, (s0, s1) in enumerate(rng_)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    re = sum(p[0] for p in parts)
    im = sum(p[1] for p in parts)
    assert np.allclose(re, a[0], rtol=1e-5, atol=1e-5 * np.abs(a[0]).max())
    assert np.allclose(im, a[1], rtol=1e-5, atol=1e-5 * np.abs(a[1]).max())
    non = abi.make_non(singles=N, tmax1=T, sample=sr, begin=0, end=ns, ts=1, propagation=1, deltat=2.0,
                       min1=lo, max1=hi)
    non.shifte = np.float32((np.float32(hi) + np.float32(lo)) / np.float32(2))
    o_re, o_im = orc.absorption(non, H, mu)
    assert np.array_equal(o_re, a[0]) and np.array_equal(o_im, a[1])
