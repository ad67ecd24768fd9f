"""Streamed absorption (nise_absorption_stream): the samples run in windows
whose frames are read from the trajectory files and uploaded while the
previous window runs, for frame ranges beyond device or host memory.

Each case is compared with the one-pass batched run on the same records
and with the oracle: windows of 7 and 5 samples with a ragged last window,
a first sample other than 0, the Full and the Coupling types, the Trotter
scheme and the exact exponential, and the window sized from the free
device memory.  Tolerances: 1e-6 relative L2 against the one-pass run
(the same per-vector FP64 sums, added per window; each window's response is
rounded to FP32 before the windows are summed in FP64), 1e-5 against the
oracle (the batched path's own bar).  A trajectory file too short for the
last window returns NISE_ERR_RANGE with a message naming it.
"""
import ctypes as C
import os

import numpy as np
import pytest

from nise_2017_amd import abi, synth

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[stream] {msg}", flush=True)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def run_stream(non, W):
    re = np.zeros(non.tmax1, np.float32)
    im = np.zeros(non.tmax1, np.float32)
    rc = abi.lib().nise_absorption_stream(C.byref(non), W, re, im)
    return rc, re, im


def files(d, H, mu, Cr=None):
    H.tofile(os.path.join(d, "Energy.bin"))
    mu.tofile(os.path.join(d, "Dipole.bin"))
    if Cr is not None:
        Cr.tofile(os.path.join(d, "Coupling.bin"))
    return (os.path.join(d, "Energy.bin").encode(), os.path.join(d, "Dipole.bin").encode(),
            os.path.join(d, "Coupling.bin").encode())


@pytest.mark.parametrize("W,prop,ctype", [(7, 1, False), (5, 1, True), (6, 0, False), (0, 1, False)])
def test_stream_matches_batched_and_oracle(orc, tmp_path, W, prop, ctype):
    N, T, sr, b, e = 12, 24, 3, 2, 31
    nfr = (e - 1) * sr + T + 5
    H, mu = synth.chain_absorption(N=N, nframes=nfr, seed=700 + W)
    Cr = None
    if ctype:
        a_, b_ = np.triu_indices(N)
        Hc = np.zeros((nfr, 1 + N), np.float32)
        Hc[:, 1:] = H[:, 1:][:, a_ == b_]
        Cr = H[0].copy()
        Cr[1:][a_ == b_] = 0
        H, mu = Hc, np.ascontiguousarray(mu[:1])
    hf, mf, cf = files(str(tmp_path), H, mu, Cr)
    lo, hi = synth.shifted_center(H, N) if not ctype else (1450.0, 1850.0)

    def mk():
        non = abi.make_non(singles=N, tmax1=T, sample=sr, begin=b, end=e, ts=1, propagation=prop, deltat=2.0,
                           min1=lo, max1=hi, hamiltonian="Coupling" if ctype else "Full", thres=0.0)
        non.shifte = np.float32((np.float32(non.max1) + np.float32(non.min1)) / np.float32(2))
        non.energyFName, non.dipoleFName, non.couplingFName = hf, mf, cf
        return non
    rc, re, im = run_stream(mk(), W)
    assert rc == 0, abi.lib().nise_last_error().decode()
    r1, i1 = np.zeros(T, np.float32), np.zeros(T, np.float32)
    abi.check(abi.lib().nise_absorption_batched(C.byref(mk()), abi.ptr(H), H.shape[1], abi.ptr(mu), mu.shape[1],
                                                H.shape[0], abi.ptr(Cr), abi.NISE_HOST, 0, r1, i1))
    o_re, o_im = orc.absorption(mk(), H, mu, C_rec=Cr)
    e1 = rel(re + 1j * im, r1 + 1j * i1)
    e2 = rel(re + 1j * im, o_re + 1j * o_im)
    say(f"window {W} propagation {prop} {'Coupling' if ctype else 'Full'}: vs one pass {e1:.1e}, vs oracle {e2:.1e}")
    assert e1 < 1e-6 and e2 < 1e-5


def test_short_file_is_an_error(tmp_path):
    N, T, sr = 8, 16, 4
    H, mu = synth.chain_absorption(N=N, nframes=60, seed=9)
    hf, mf, _ = files(str(tmp_path), H[:50], mu)
    non = abi.make_non(singles=N, tmax1=T, sample=sr, begin=0, end=10, ts=1, propagation=1, deltat=2.0,
                       min1=1550.0, max1=1750.0)
    non.energyFName, non.dipoleFName = hf, mf
    rc, _, _ = run_stream(non, 3)
    msg = abi.lib().nise_last_error().decode()
    say(f"rc {rc}: {msg}")
    assert rc == abi.NISE_ERR_RANGE and "Energy.bin" in msg
