"""The GPU 2D drop-ins against the reference's own calc_2DIR / calc_2DES.

Every case of tests/golden/reference_2d.npz (written by the reference's
calc_2DIR.c / calc_2DES.c compiled unmodified and run as an MPI singleton:
coupled rings and a chain, per-frame anharmonicities, Trotter 2, each
pathway, a Coupling-type Hamiltonian, 2DUVvis, waiting times t2 > 0 and the
default sparse propagation) is run through the library's calc_2DIR /
calc_2DES with the same input files and t_non, and the six R*.dat files are
compared with the reference's.

Beyond those small decks, nise_2dir_batched is compared with the oracle
(which the CPU tests pin to the same reference files) on larger coupled
systems with a waiting time and with the sparse scheme.

Tolerances (relative L2 over the six files' response columns): 1e-5 for
the Trotter scheme (FP32 propagation in the reference's order; the sums
over items and samples are FP64 here, FP32 running sums there, and the
files carry 7 digits); 1e-4 where a frame is diagonalised (the reference's
FP32 ssyev_ against the FP64 Jacobi solver here, magnified by the
cancellation of the GB, SE and EA pathways; the oracle shows the same
2e-5).
"""
import ctypes as C
import os

import numpy as np
import pytest

from nise_2017_amd import abi, synth

import ref2d_cases as RC

pytestmark = pytest.mark.gpu

DIAG_CASES = {"ring6_t2", "ring5_2duvvis_t2", "ring6_sparse", "ring5_sparse_t2", "ring5_sparse_2duvvis"}


def say(msg):
    print(f"[ref2d] {msg}", flush=True)


def _dropin(name):
    L = abi.lib()
    f = getattr(L, name)
    f.argtypes = [C.POINTER(abi.TNon), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    f.restype = None
    return f


@pytest.mark.parametrize("name", RC.names())
def test_dropin_against_reference_files(name, tmp_path, monkeypatch):
    non, rec, deck, R = RC.case(name)
    d = str(tmp_path)
    fields = {"Energy": "energyFName", "Dipole": "dipoleFName", "Anharmonicity": "anharFName",
              "OvertoneDipole": "overdipFName", "Coupling": "couplingFName"}
    for key, arr in rec.items():
        path = os.path.join(d, key + ".bin")
        arr.tofile(path)
        setattr(non, fields[key], path.encode())
    say(f"{name}: {deck['Technique']}, propagation {non.propagation}, t2 {non.tmax2}")
    monkeypatch.chdir(d)
    fn = _dropin("calc_2DES" if deck["Technique"].endswith("UVvis") else "calc_2DIR")
    fn(C.byref(non), 0, 1, 0, 1, 0, 0)
    got = np.stack([np.loadtxt(os.path.join(d, f), ndmin=2) for f in RC.FILES])
    assert got.shape == R.shape
    assert np.allclose(got[..., :3], R[..., :3])
    e = RC.rel(got[..., 3:], R[..., 3:])
    say(f"{name}: rel L2 {e:.2e}")
    assert e < (1e-4 if name in DIAG_CASES else 1e-5)


def _system(N, T1, T2, sr, ns, seed, ring=True):
    nfr = (ns - 1) * sr + T1 + T2 + T1 + 3
    H, mu, anh, over = synth.twodir_system(N=N, nframes=nfr, seed=seed, sd=12.0, jmean=-6.0, dip_fluct=0.15)
    if ring:
        a, b = np.triu_indices(N)
        H[:, 1 + np.nonzero((a == 0) & (b == N - 1))[0][0]] = np.float32(-5.5)
    return H, mu, anh, over


@pytest.mark.parametrize("N,T1,T2,prop,tech,ts,thr", [
    (12, 10, 7, 1, "2DIR", 1, 0.0),       # waiting time, Trotter scheme (persistent EA kernel)
    (9, 8, 5, 1, "2DUVvis", 2, 0.0),      # waiting time, two-level sites, Trotter 2
    (10, 8, 0, 0, "2DIR", 1, 0.001),      # the default sparse scheme
    (8, 6, 4, 0, "2DIR", 2, 0.01),        # sparse scheme with a waiting time, Trotter 2
    (7, 6, 0, 0, "2DUVvis", 1, 0.0),      # exact exponential (threshold 0), two-level sites
    (6, 6, 3, 0, "noEAIR", 1, 0.001),
])
def test_batched_against_oracle(orc, N, T1, T2, prop, tech, ts, thr):
    sr, ns = 4, 5
    H, mu, anh, over = _system(N, T1, T2, sr, ns, seed=200 + N + T2)
    lo, hi = synth.shifted_center(H, N)
    g = np.float32(np.float32(np.float32(np.float32(2.0) * RC.ICM2IFS) * RC.TWOPI) / np.float32(ts))

    def mk():
        non = abi.make_non(singles=N, tmax1=T1, tmax2=T2, tmax3=T1, sample=sr, begin=0, end=ns, ts=ts,
                           propagation=prop, deltat=2.0, min1=lo, max1=hi, technique=tech,
                           anharmonicity=16.0, lifetime=500.0)
        non.thres = np.float32(np.float32(thr) * np.float32(g * g))
        return non
    say(f"N={N} T1={T1} T2={T2} propagation {prop} {tech} Trotter {ts} threshold {thr}")
    R = abi.twodir_batched(mk(), H, mu)
    ref = orc.twodir(mk(), H, mu, acc64=True)
    e = RC.rel(R, ref)
    say(f"rel L2 {e:.2e}")
    assert e < 1e-4
