"""Pins of the 2DFFT oracle (oracle/fft2d.py, a restatement of
src/2DFFT.c:20-830).  The reference program needs FFTW, which this image
does not have, so it is not built here; the oracle is pinned by:

- numpy's forward transform against the brute-force DFT sum with the sign
  and layout of FFTW_FORWARD;
- a delta response at t1 = t3 = 0: the two edge halvings leave 1/4 on every
  grid point of both transforms, and 2D = kI + kII = 1/2 at every written
  point, PP = p1 / 2 (the imaginary column);
- a single-frequency non-rephasing response exp(-2 pi i (m1 t1 + m3 t3) /
  fft): its only peak sits at w1 = shift1 - m1 dw, w3 = shift3 - m3 dw
  (dw = 1 / (deltat c fft)) with height (tmax1 - 1/2)(tmax3 - 1/2).
"""
import os

import numpy as np
import pytest

from nise_2017_amd import abi
from oracle import fft2d

CV = 2.99792458e-5


def write_print2d(path, R, dt):
    """print2D layout (src/MPI_subs.c:20-34): t1 outer, t3 inner."""
    T1, T3 = R.shape
    with open(path, "w") as fh:
        for t1 in range(T1):
            for t3 in range(T3):
                fh.write("%f %f %f %e %e\n" % (t1 * dt, 0.0, t3 * dt, R[t1, t3].real, R[t1, t3].imag))


def load(path):
    return np.loadtxt(path, ndmin=2)


def test_numpy_forward_transform_is_the_dft_sum():
    rng = np.random.default_rng(3)
    n = 8
    x = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    k = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(k, k) / n)
    assert np.allclose(np.fft.fft2(x), W @ x @ W.T, rtol=1e-12, atol=1e-12)


def mk(T, fft_, **kw):
    base = dict(tmax1=T, tmax2=0, tmax3=T, fft=fft_, deltat=2.0, min1=1000.0, max1=3000.0, min2=0.0,
                max2=0.0, min3=1000.0, max3=3000.0, technique="2DIR")
    base.update(kw)
    return abi.make_non(**base)


def test_delta_response(tmp_path):
    T, fft = 6, 64
    non = mk(T, fft)
    R = np.zeros((T, T), complex)
    R[0, 0] = 1 + 1j
    for name in ("RparI", "RparII", "RperI", "RperII", "RcroI", "RcroII"):
        write_print2d(tmp_path / f"{name}.dat", R, non.deltat)
    assert fft2d.run(non, 0, str(tmp_path))
    for pol in ("par", "per", "cro"):
        for k in ("I", "II"):
            w = load(tmp_path / f"Rw{pol}.{k}.dat")
            assert len(w) > 4 and np.all(w[:, 2] == 0.25) and np.all(w[:, 3] == 0.25)
        d = load(tmp_path / f"2D.{pol}.dat")
        assert np.all(d[:, 2] == 0.5) and np.all(d[:, 3] == 0.5)
        pp = np.loadtxt(tmp_path / f"PP.{pol}.dat", comments="#", ndmin=2)
        p1 = len(d) // len(pp)
        assert len(d) == p1 * len(pp) and np.all(pp[:, 1] == 0.5 * p1)


@pytest.mark.parametrize("m1,m3", [(2, 3), (3, 1)])
def test_single_frequency_peak(tmp_path, m1, m3):
    T, fft = 8, 64
    non = mk(T, fft)
    t = np.arange(T)
    R = np.exp(-2j * np.pi * (m1 * t[:, None] + m3 * t[None, :]) / fft)
    for pol in ("par", "per", "cro"):
        write_print2d(tmp_path / f"R{pol}I.dat", np.zeros((T, T), complex), non.deltat)
        write_print2d(tmp_path / f"R{pol}II.dat", R, non.deltat)
    assert fft2d.run(non, 0, str(tmp_path))
    w = load(tmp_path / "Rwpar.II.dat")
    mag = np.hypot(w[:, 2], w[:, 3])
    k = int(np.argmax(mag))
    dw = 1 / (non.deltat * CV * fft)
    shift1 = 0.5 * (non.min1 + non.max1)
    shift3 = 0.5 * (non.min3 + non.max3)
    assert abs(w[k, 0] - (shift1 - m1 * dw)) < 1e-3
    assert abs(w[k, 1] - (shift3 - m3 * dw)) < 1e-3
    assert abs(w[k, 2] - (T - 0.5) ** 2) < 1e-5 * (T - 0.5) ** 2 and abs(w[k, 3]) < 1e-5
    # kI is zero, so the sum file repeats the kII file point by point
    d = load(tmp_path / "2D.par.dat")
    assert d.shape == w.shape and np.allclose(d, w, rtol=1e-6, atol=1e-6)


def test_bad_ranges_exit(tmp_path):
    assert not fft2d.run(mk(4, 32, min1=-1.0), 0, str(tmp_path))
    assert not fft2d.run(mk(4, 32, min3=4000.0), 0, str(tmp_path))
    assert not fft2d.run(mk(4, 0), 0, str(tmp_path))
