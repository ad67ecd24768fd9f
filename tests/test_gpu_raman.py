"""GPU parity of 2DIR-Raman (nise_2dirraman_batched and the calc_2DIRraman
file drop-in) against the oracle's or_2dirraman_item, which the CPU tests
pin to the reference's own primitives (dipole_double_ground,
find_raman_pol, read_alpha) and to the exact response of uncoupled sites.

Coupled rings with fluctuating energies, dipoles and Raman tensors, every
technique variant, Trotter 1 and 2, the Full and the Coupling Hamiltonian
types, a waiting time in the input (which the reference's Raman driver
reads only for the sample count).  Tolerance 1e-5 relative L2 over the 12
planes: the propagation follows the reference's FP32 operations (the
doubles step's cos/sin rounded once from double, within an ulp of the
reference's), each item's terms are the reference's FP32 values, and the
sums over items and samples are FP64 here.
"""
import ctypes as C
import os

import numpy as np
import pytest

from nise_2017_amd import abi, synth

import ref2d_cases as RC

pytestmark = pytest.mark.gpu
TECHS = ["2DIRraman", "2DIRraman1", "2DIRraman2", "2DIRraman3", "2DIRramanI", "2DIRramanII"]


def say(msg):
    print(f"[raman] {msg}", flush=True)


def system(N, nfr, seed, ring=True):
    H, mu, _, _ = synth.twodir_system(N=N, nframes=nfr, seed=seed, sd=12.0, jmean=-6.0, dip_fluct=0.15)
    if ring and N > 2:
        a, b = np.triu_indices(N)
        H[:, 1 + np.nonzero((a == 0) & (b == N - 1))[0][0]] = np.float32(-5.5)
    rng = np.random.default_rng(seed + 1)
    al = np.zeros((nfr, 1 + 6 * N), np.float32)
    al[:, 1:] = (0.6 + 0.3 * rng.normal(size=(nfr, 6 * N))).astype(np.float32)
    return H, mu, al


def make(N, T1, T3, sr, ns, tech, ts=1, T2=0, H=None, ctype=False):
    lo, hi = synth.shifted_center(H, N) if H is not None else (1550.0, 1750.0)
    return abi.make_non(singles=N, tmax1=T1, tmax2=T2, tmax3=T3, sample=sr, begin=0, end=ns, ts=ts,
                        propagation=1, deltat=2.0, min1=lo, max1=hi, technique=tech, anharmonicity=16.0,
                        lifetime=400.0, hamiltonian="Coupling" if ctype else "Full")


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("tech", TECHS)
def test_batched_against_oracle(orc, tech):
    N, T1, T3, sr, ns = 7, 5, 6, 3, 4
    nfr = (ns - 1) * sr + T1 + T3 + 2
    H, mu, al = system(N, nfr, seed=300 + TECHS.index(tech))
    say(f"{tech}: N={N} T1={T1} T3={T3} {ns} samples")
    R = abi.twodir_raman_batched(make(N, T1, T3, sr, ns, tech, H=H), H, mu, al)
    ref = orc.twodir_raman(make(N, T1, T3, sr, ns, tech, H=H), H, mu, al, acc64=True)
    e = rel(R, ref)
    say(f"{tech}: rel L2 {e:.2e}")
    assert e < 1e-5


@pytest.mark.parametrize("N,T1,T3,ts,T2", [(12, 8, 7, 2, 0), (5, 1, 4, 1, 3), (9, 6, 1, 1, 0), (40, 5, 4, 1, 0)])
def test_shapes_against_oracle(orc, N, T1, T3, ts, T2):
    """Trotter 2, a waiting time in the input, t1 or t3 of one step, and
    N = 40: one kII vector per wave (64 lanes per vector), an odd t1 count."""
    sr, ns = 4, 3
    nfr = (ns - 1) * sr + T1 + T2 + T3 + 2
    H, mu, al = system(N, nfr, seed=400 + N)
    say(f"N={N} T1={T1} T3={T3} Trotter {ts} t2 {T2}")
    R = abi.twodir_raman_batched(make(N, T1, T3, sr, ns, "2DIRraman", ts=ts, T2=T2, H=H), H, mu, al)
    ref = orc.twodir_raman(make(N, T1, T3, sr, ns, "2DIRraman", ts=ts, T2=T2, H=H), H, mu, al, acc64=True)
    e = rel(R, ref)
    say(f"rel L2 {e:.2e}")
    assert e < 1e-5


def coupling_records(H, mu, N):
    """Coupling type: per-frame energies, the couplings of frame 0, static dipoles."""
    a, b = np.triu_indices(N)
    Hc = np.zeros((H.shape[0], 1 + N), np.float32)
    Hc[:, 1:] = H[:, 1:][:, a == b]
    Cr = H[0].copy()
    Cr[1:][a == b] = 0
    return Hc, np.ascontiguousarray(mu[:1]), Cr


def test_coupling_type_and_device_records(orc):
    import torch
    N, T1, T3, sr, ns = 8, 4, 5, 3, 3
    nfr = (ns - 1) * sr + T1 + T3 + 2
    H, mu, al = system(N, nfr, seed=500)
    Hc, muc, Cr = coupling_records(H, mu, N)
    mk = lambda: make(N, T1, T3, sr, ns, "2DIRraman", H=H, ctype=True)  # noqa: E731
    R = abi.twodir_raman_batched(mk(), Hc, muc, al, coupling_rec=Cr)
    ref = orc.twodir_raman(mk(), Hc, muc, al, C_rec=Cr, acc64=True)
    e = rel(R, ref)
    say(f"Coupling type: rel L2 {e:.2e}")
    assert e < 1e-5
    Rd = abi.twodir_raman_batched(mk(), torch.from_numpy(Hc).cuda(), torch.from_numpy(muc).cuda(),
                                  torch.from_numpy(al).cuda(), coupling_rec=torch.from_numpy(Cr).cuda())
    assert np.array_equal(R, Rd)


def test_errors():
    N, T1, T3, sr, ns = 4, 3, 3, 2, 2
    nfr = 20
    H, mu, al = system(N, nfr, seed=9)
    for kw, code in ((dict(propagation=0), abi.NISE_ERR_ARG), (dict(anharmonicity=0.0), abi.NISE_ERR_ARG),
                     (dict(technique="2DIR"), abi.NISE_ERR_ARG), (dict(end=50), abi.NISE_ERR_RANGE)):
        non = make(N, T1, T3, sr, ns, "2DIRraman", H=H)
        for k, v in kw.items():
            setattr(non, k, v.encode() if isinstance(v, str) else v)
        R = np.zeros((12, T3, T1), np.float32)
        rc = abi.lib().nise_2dirraman_batched(C.byref(non), abi.ptr(H), H.shape[1], abi.ptr(mu), mu.shape[1],
                                              nfr, None, abi.ptr(al), abi.NISE_HOST, R)
        say(f"{kw}: rc {rc} ({abi.lib().nise_last_error().decode()})")
        assert rc == code


def test_dropin_files(orc, tmp_path, monkeypatch):
    N, T1, T3, sr, ns = 6, 4, 5, 3, 3
    nfr = (ns - 1) * sr + T1 + T3 + 2
    H, mu, al = system(N, nfr, seed=600)
    d = str(tmp_path)
    for name, arr in (("Energy.bin", H), ("Dipole.bin", mu), ("Alpha.bin", al)):
        arr.tofile(os.path.join(d, name))
    non = make(N, T1, T3, sr, ns, "2DIRraman", H=H)
    non.energyFName, non.dipoleFName, non.alphaFName = b"Energy.bin", b"Dipole.bin", b"Alpha.bin"
    non.length = nfr
    monkeypatch.chdir(d)
    f = abi.lib().calc_2DIRraman
    f.argtypes = [C.POINTER(abi.TNon), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    f.restype = None
    f(C.byref(non), 0, 1, 0, 1, 0, 0)
    names = ["RparI", "RparII", "RperI", "RperII", "RcroI", "RcroII"]
    got = np.stack([np.loadtxt(os.path.join(d, n + ".IRraman.dat"), ndmin=2) for n in names])
    ref = RC.print2d(non, orc.twodir_raman(make(N, T1, T3, sr, ns, "2DIRraman", H=H), H, mu, al), ns)
    assert got.shape == ref.shape
    assert np.allclose(got[..., :3], ref[..., :3])
    e = RC.rel(got[..., 3:], ref[..., 3:])
    say(f"drop-in files: rel L2 {e:.2e}")
    assert e < 1e-5
