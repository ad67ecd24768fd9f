"""GPU parity of couplings built per frame from geometry
(HamiltonianType TransitionDipole / ExtendedDipole,
src/read_trajectory.c:181-303) through the C ABI: the on-the-fly plan
(nise_abs_plan_create_onfly) and the file drop-in absorption() reading the
position and pbc files, against the oracle on the same inputs.

The couplings are evaluated in IEEE FP32 in the reference's source order on
both sides (the oracle's restatement is bit-identical to the reference's
read_He built with its ACCURATE_MATHS flags); the propagation is the
reference's operation order; responses are summed in FP64 here and FP32 in
the oracle, hence 1e-5 relative L2 on the response."""
import ctypes as C
import os

import numpy as np
import pytest

from nise_2017_amd import abi, synth
from nise_2017_amd.abi import make_non

pytestmark = pytest.mark.gpu

HTYPE = {1: "TransitionDipole", 2: "ExtendedDipole"}


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


def oracle_absorption(orc, non, H, mu, pos, box):
    orc.oracle().or_set_geometry(orc.ptr(pos), pos.shape[1], orc.ptr(mu), mu.shape[1], orc.ptr(box))
    try:
        return orc.absorption(non, H, mu)
    finally:
        orc.oracle().or_set_geometry(None, 0, None, 0, None)


@pytest.mark.parametrize("kind,box", [(1, None), (1, 11.0), (2, None), (2, 11.0)])
def test_onfly_plan_vs_oracle(orc, kind, box):
    N, T, sr, ns = 12, 40, 7, 6
    nfr = (ns - 1) * sr + T + 3
    H, mu, pos, boxes = synth.onfly_system(N, nfr, 60 + kind, kind=kind, box=box)
    lo, hi = 1550.0, 1750.0
    non = make_non(singles=N, tmax1=T, sample=sr, begin=0, end=ns, ts=1, propagation=1, deltat=2.0, min1=lo,
                   max1=hi, hamiltonian=HTYPE[kind], couplingcut=0.02)
    non.shifte = np.float32((np.float32(hi) + np.float32(lo)) / np.float32(2))
    p = abi.AbsPlan(non, H, H.shape[1], mu, mu.shape[1], nfr, positions=pos, box=boxes)
    re, im = p.execute()
    p.close()
    o_re, o_im = oracle_absorption(orc, non, H, mu, pos, boxes)
    assert rel(re + 1j * im, o_re + 1j * o_im) < 1e-5
    # the geometry matters: without the box (or with the other scheme) the answer moves
    if box is not None:
        n_re, n_im = oracle_absorption(orc, non, H, mu, pos, None)
        assert rel(n_re + 1j * n_im, o_re + 1j * o_im) > 1e-4


def test_onfly_without_positions_is_an_error(orc):
    N, nfr = 4, 30
    H, mu, pos, _ = synth.onfly_system(N, nfr, 5, kind=1)
    non = make_non(singles=N, tmax1=8, sample=5, begin=0, end=2, ts=1, propagation=1, deltat=2.0,
                   min1=1550.0, max1=1750.0, hamiltonian="TransitionDipole")
    with pytest.raises(RuntimeError):
        abi.AbsPlan(non, H, H.shape[1], mu, mu.shape[1], nfr)


@pytest.fixture(scope="module")
def L():
    L = abi.lib()
    L.absorption.argtypes = [C.POINTER(abi.TNon)]
    L.absorption.restype = None
    return L


@pytest.mark.parametrize("kind,box", [(1, 11.0), (2, None)])
def test_absorption_dropin_reads_geometry(orc, L, tmp_path, monkeypatch, kind, box):
    N, T, sr, ns = 10, 32, 9, 5
    nfr = (ns - 1) * sr + T + 4
    H, mu, pos, boxes = synth.onfly_system(N, nfr, 80 + kind, kind=kind, box=box)
    d = str(tmp_path)
    files = dict(energyFName="H.bin", dipoleFName="mu.bin", positionFName="pos.bin")
    H.tofile(os.path.join(d, "H.bin"))
    mu.tofile(os.path.join(d, "mu.bin"))
    pos.tofile(os.path.join(d, "pos.bin"))
    if boxes is not None:
        boxes.tofile(os.path.join(d, "box.bin"))
        files["pbcFName"] = "box.bin"
    kw = dict(singles=N, tmax1=T, sample=sr, length=nfr, ts=1, propagation=1, deltat=2.0, min1=1550.0,
              max1=1750.0, fft=256, lifetime=300.0, end=ns, hamiltonian=HTYPE[kind], couplingcut=0.02)
    non = make_non(**{k: os.path.join(d, v) for k, v in files.items()}, **kw)
    monkeypatch.chdir(tmp_path)
    L.absorption(C.byref(non))
    td = np.loadtxt("TD_Absorption.dat", ndmin=2)
    onon = make_non(**kw)
    onon.shifte = non.shifte
    re, im = oracle_absorption(orc, onon, H, mu, pos, boxes)
    assert rel(td[:, 1] + 1j * td[:, 2], (re + 1j * im) / ns) < 1e-5
