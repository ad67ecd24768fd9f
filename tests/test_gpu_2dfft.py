"""GPU parity of the in-process 2DFFT (nise_2dfft, src/2DFFT.c:20-830) and
its device transform (nise_dft2d), through the C ABI.

nise_dft2d is checked against numpy's forward FFT (FP64, 1e-12 relative).
nise_2dfft runs on response files in one directory and the oracle
(oracle/fft2d.py) on the same files in another; every output file must
exist in both, axis columns must be identical (the same FP32 arithmetic in
source order), and value columns must agree to 2e-6 of the file's largest
value: the transforms differ by rounding (~1e-15) and the files carry %e
(7 significant digits), so a value can differ in its last printed digit,
also after the sums re-read those digits.
"""
import ctypes as C
import os
import shutil

import numpy as np
import pytest

from nise_2017_amd import abi
from oracle import fft2d

pytestmark = pytest.mark.gpu

CV = 2.99792458e-5
POLS = ("par", "per", "cro")


def say(msg):
    print(f"[2dfft] {msg}", flush=True)


def test_dft2d_matches_numpy():
    say("nise_dft2d vs numpy fft2")
    rng = np.random.default_rng(11)
    P, fft, nr, nc = 3, 64, 20, 30
    x = np.zeros((P, fft, fft), np.complex128)
    x[:, :nr, :nc] = rng.normal(size=(P, nr, nc)) + 1j * rng.normal(size=(P, nr, nc))
    out = np.zeros_like(x)
    abi.check(abi.lib().nise_dft2d(x.ctypes.data_as(C.c_void_p), P, fft, nr, nc, out.ctypes.data_as(C.c_void_p)))
    ref = np.fft.fft2(x)
    assert np.max(np.abs(out - ref)) < 1e-12 * np.max(np.abs(ref))


def responses(T1, T3, dt, seed):
    """Three decaying oscillations per response, in the rotating frame."""
    rng = np.random.default_rng(seed)
    t1 = np.arange(T1)[:, None] * dt
    t3 = np.arange(T3)[None, :] * dt
    R = np.zeros((T1, T3), complex)
    for _ in range(3):
        a = rng.normal() + 1j * rng.normal()
        f1, f3 = rng.uniform(-120, 120, 2) * 2 * np.pi * CV
        R += a * np.exp(-1j * (f1 * t1 + f3 * t3) - (t1 + t3) / 600.0)
    return R


def write_print2d(path, R, dt):
    T1, T3 = R.shape
    with open(path, "w") as fh:
        for a in range(T1):
            for b in range(T3):
                fh.write("%f %f %f %e %e\n" % (a * dt, 0.0, b * dt, R[a, b].real, R[a, b].imag))


def load(path):
    rows = []
    with open(path) as fh:
        for line in fh:
            if line.startswith("#") or not line.strip():
                continue
            rows.append([float(v) for v in line.split()])
    return rows


def compare(dir_a, dir_b, naxis_of):
    fa, fb = sorted(os.listdir(dir_a)), sorted(os.listdir(dir_b))
    assert fa == fb, (fa, fb)
    identical = 0
    for name in fa:
        ta, tb = open(os.path.join(dir_a, name)).read(), open(os.path.join(dir_b, name)).read()
        if ta == tb:
            identical += 1
            continue
        A, B = load(os.path.join(dir_a, name)), load(os.path.join(dir_b, name))
        assert len(A) == len(B), name
        nax = naxis_of(name)
        scale = max((abs(v) for r in B for v in r[nax:]), default=0.0)
        for ra, rb in zip(A, B):
            assert len(ra) == len(rb), name
            assert ra[:nax] == rb[:nax], (name, ra, rb)
            for x, y in zip(ra[nax:], rb[nax:]):
                assert abs(x - y) <= 2e-6 * scale, (name, ra, rb)
    return identical, len(fa)


def naxis(form):
    def f(name):
        if name.startswith("R") and not name.startswith("Rw"):
            return 3  # inputs
        if name == "waxis.dat":
            return 1
        if name.startswith("PP"):
            return 1
        return 0 if form == 1 else 2
    return f


def run_both(tmp_path, monkeypatch, non_kw, form, files):
    a, b = tmp_path / "gpu", tmp_path / "oracle"
    a.mkdir(), b.mkdir()
    for name, text in files.items():
        (a / name).write_text(text)
        (b / name).write_text(text)
    non = abi.make_non(**non_kw)
    monkeypatch.chdir(a)
    rc = abi.lib().nise_2dfft(C.byref(non), form)
    monkeypatch.chdir(tmp_path)
    assert fft2d.run(abi.make_non(**non_kw), form, str(b))
    return rc, str(a), str(b)


def print2d_text(R, dt):
    T1, T3 = R.shape
    return "".join("%f %f %f %e %e\n" % (x * dt, 0.0, y * dt, R[x, y].real, R[x, y].imag)
                   for x in range(T1) for y in range(T3))


def base_kw(**kw):
    d = dict(tmax1=16, tmax2=0, tmax3=12, fft=128, deltat=2.0, min1=1500.0, max1=1800.0, min2=0.0,
             max2=0.0, min3=1450.0, max3=1850.0, technique="2DIR")
    d.update(kw)
    return d


@pytest.mark.parametrize("form,homo,inhomo,window", [(0, 0.0, 0.0, 0), (2, 300.0, 0.0, 0),
                                                     (0, 0.0, 150.0, 1), (1, 0.0, 0.0, 0)])
def test_2dfft_files_vs_oracle(tmp_path, monkeypatch, form, homo, inhomo, window):
    say(f"2dfft files format={form} homogen={homo} inhomogen={inhomo} window={window}")
    kw = base_kw(homogen=homo, inhomogen=inhomo, window=window)
    files = {}
    for i, pol in enumerate(POLS):
        for k in ("I", "II"):
            files[f"R{pol}{k}.dat"] = print2d_text(responses(16, 12, 2.0, 7 * i + len(k)), 2.0)
    rc, a, b = run_both(tmp_path, monkeypatch, kw, form, files)
    assert rc == abi.NISE_OK
    same, n = compare(a, b, naxis(form))
    say(f"{same} of {n} files byte-identical")
    names = os.listdir(a)
    assert "Rwpar.I.dat" in names and "Rwcro.II.dat" in names
    if form != 1:
        assert "2D.per.dat" in names and "PP.cro.dat" in names


def test_2dfft_waiting_times_and_missing_pol(tmp_path, monkeypatch):
    say("2dfft waiting times, per polarisation missing")
    files = {"Waitingtime.dat": "100\n200\n"}
    for wt in ("100", "200"):
        for i, pol in enumerate(("par", "cro")):
            for k in ("I", "II"):
                files[f"R{pol}{k}_{wt}fs.dat"] = print2d_text(responses(16, 12, 2.0, 31 * i + int(wt)), 2.0)
    rc, a, b = run_both(tmp_path, monkeypatch, base_kw(), 0, files)
    assert rc == abi.NISE_OK
    compare(a, b, naxis(0))
    names = os.listdir(a)
    assert "2D.par_100fs.dat" in names and "PP.cro_200fs.dat" in names and "2D.per_100fs.dat" not in names


def test_2dfft_raman_names(tmp_path, monkeypatch):
    say("2dfft 2DIRraman names")
    files = {}
    for i, pol in enumerate(POLS):
        for k in ("I", "II"):
            files[f"R{pol}{k}.IRraman.dat"] = print2d_text(responses(16, 12, 2.0, 5 * i + len(k)), 2.0)
    rc, a, b = run_both(tmp_path, monkeypatch, base_kw(technique="2DIRraman2"), 0, files)
    assert rc == abi.NISE_OK
    compare(a, b, naxis(0))
    names = os.listdir(a)
    for n in ("2DIRraman.I.par.dat", "2DIRraman.II.cro.dat", "PPIRraman.II.per.dat", "RwperIRraman.I.dat"):
        assert n in names, n


def test_2dfft_after_calc_2dir(orc, tmp_path, monkeypatch):
    """The files calc_2DIR writes, transformed in process."""
    from nise_2017_amd import synth
    say("calc_2DIR -> nise_2dfft")
    N, T1, T3, sr = 4, 8, 8, 5
    H, mu, anh, over = synth.twodir_system(N=N, nframes=60, seed=61)
    lo, hi = synth.shifted_center(H, N)
    d = str(tmp_path / "run")
    os.makedirs(d)
    H.tofile(os.path.join(d, "H.bin")), mu.tofile(os.path.join(d, "mu.bin"))
    kw = dict(singles=N, tmax1=T1, tmax2=0, tmax3=T3, sample=sr, begin=0, end=4, ts=1, propagation=1,
              deltat=2.0, min1=lo, max1=hi, min3=lo, max3=hi, couplingcut=0.0, technique="2DIR",
              anharmonicity=16.0, lifetime=0.0, length=60, fft=64)
    L = abi.lib()
    L.calc_2DIR.argtypes = [C.POINTER(abi.TNon), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    L.calc_2DIR.restype = None
    non = abi.make_non(energyFName=os.path.join(d, "H.bin"), dipoleFName=os.path.join(d, "mu.bin"), **kw)
    monkeypatch.chdir(d)
    L.calc_2DIR(C.byref(non), 0, 1, 0, 1, 0, 0)
    b = str(tmp_path / "oracle")
    os.makedirs(b)
    for name in os.listdir(d):
        if name.startswith("R") and name.endswith(".dat"):
            shutil.copy(os.path.join(d, name), b)
    assert L.nise_2dfft(C.byref(abi.make_non(**kw)), 0) == abi.NISE_OK
    assert fft2d.run(abi.make_non(**kw), 0, b)
    for name in ("H.bin", "mu.bin"):
        os.remove(os.path.join(d, name))
    compare(d, b, naxis(0))


def test_2dfft_errors(tmp_path, monkeypatch):
    say("2dfft error codes")
    monkeypatch.chdir(tmp_path)
    L = abi.lib()
    for kw in (dict(min1=-5.0), dict(min3=1900.0), dict(fft=0)):
        assert L.nise_2dfft(C.byref(abi.make_non(**base_kw(**kw))), 0) == abi.NISE_ERR_ARG
