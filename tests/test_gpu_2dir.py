"""GPU parity of the batched 2DIR response (nise_2dir_batched) and the
calc_2DIR file drop-in.

Oracle: oracle.twodir (the reference's per-item loop, src/calc_2DIR.c
:184-506, summed over samples and the 21 polarisation items in work order)
and, for uncoupled sites, the closed form of twodir_closed.py.  Tolerance:
1e-4 relative L2 over the 12 planes (FP32 propagation is bit-exact per
vector; the sums over samples, items and pathways are in a different order
and accumulated in FP64 here).
"""
import ctypes as C
import os

import numpy as np
import pytest

from nise_2017_amd import abi, synth

import twodir_closed as CF

pytestmark = pytest.mark.gpu

TOL = 1e-4


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def mk(N, T1, T3, sr, begin, end, H, **kw):
    lo, hi = synth.shifted_center(H, N)
    base = dict(singles=N, tmax1=T1, tmax2=0, tmax3=T3, sample=sr, begin=begin, end=end, ts=1,
                propagation=1, deltat=2.0, min1=lo, max1=hi, couplingcut=0.0, technique="2DIR",
                anharmonicity=16.0, lifetime=0.0)
    base.update(kw)
    return abi.make_non(**base)


def say(msg):
    print(f"[2dir] {msg}", flush=True)


@pytest.mark.parametrize("N,T1,T3,ts,tech", [(6, 8, 8, 1, "2DIR"), (5, 6, 9, 2, "2DIR"),
                                              (4, 5, 5, 1, "GBIR"), (4, 5, 5, 1, "SEIR"),
                                              (4, 5, 5, 1, "EAIR"), (4, 5, 5, 1, "noEAIR")])
def test_chain_vs_oracle(orc, N, T1, T3, ts, tech):
    say(f"chain N={N} T1={T1} T3={T3} ts={ts} {tech}")
    H, mu, anh, over = synth.twodir_system(N=N, nframes=80, seed=31 + N)
    non = mk(N, T1, T3, 7, 1, 5, H, ts=ts, technique=tech, lifetime=400.0)
    ref = orc.twodir(non, H, mu)
    non = mk(N, T1, T3, 7, 1, 5, H, ts=ts, technique=tech, lifetime=400.0)
    R = abi.twodir_batched(non, H, mu)
    assert rel(R, ref) < TOL


def test_anharmonicity_records_vs_oracle(orc):
    say("anharmonicity records N=5")
    N, T1, T3 = 5, 6, 6
    H, mu, anh, over = synth.twodir_system(N=N, nframes=70, seed=41)
    non = mk(N, T1, T3, 5, 0, 6, H, anharmonicity=0.0)
    ref = orc.twodir(non, H, mu, A_rec=anh, over_rec=over)
    non = mk(N, T1, T3, 5, 0, 6, H, anharmonicity=0.0)
    R = abi.twodir_batched(non, H, mu, anh_rec=anh, over_rec=over)
    assert rel(R, ref) < TOL


def test_coupling_type_vs_oracle(orc):
    say("Coupling type N=9 lattice")
    N, T1, T3, nfr = 9, 6, 6, 60
    rng, mus, coup = synth.lattice_static(3, seed=43, cut=0.1)
    Hf = synth.full_records(synth.lattice_diag(rng, nfr, N), coup)
    a, b = np.triu_indices(N)
    Hc = np.ascontiguousarray(np.concatenate([Hf[:, :1], Hf[:, 1:][:, a == b]], axis=1))
    Cr = np.ascontiguousarray(Hf[0])
    mur = synth.dipole_records(mus, 1)
    kw = dict(hamiltonian="Coupling", couplingcut=0.5)
    non = mk(N, T1, T3, 6, 0, 5, Hf, **kw)
    ref = orc.twodir(non, Hc, mur, C_rec=Cr)
    non = mk(N, T1, T3, 6, 0, 5, Hf, **kw)
    R = abi.twodir_batched(non, Hc, mur, coupling_rec=Cr)
    assert rel(R, ref) < TOL


def test_uncoupled_closed_form():
    say("uncoupled closed form N=3")
    N, T1, T3 = 3, 7, 5
    H, mu, anh, over = synth.twodir_system(N=N, nframes=60, coupled=False, seed=44)
    non = mk(N, T1, T3, 4, 2, 6, H, lifetime=250.0)
    R = abi.twodir_batched(non, H, mu)
    assert rel(R, CF.two_level(non, H, mu)) < TOL


def test_larger_system_vs_oracle(orc):
    """N = 70: doubles vectors wider than one wave, many rotation layers."""
    say("N=70 chain")
    N, T1, T3 = 70, 4, 4
    H, mu, anh, over = synth.twodir_system(N=N, nframes=40, seed=45)
    non = mk(N, T1, T3, 9, 0, 2, H)
    ref = orc.twodir(non, H, mu)
    non = mk(N, T1, T3, 9, 0, 2, H)
    R = abi.twodir_batched(non, H, mu)
    assert rel(R, ref) < TOL


def test_partial_t1_chunk_vs_oracle(orc):
    """N = 32, T1 = 13: the EA kernel stages 8 doubles vectors per workgroup,
    so the second t1 chunk is partial."""
    say("N=32 T1=13 partial EA chunk")
    N, T1, T3 = 32, 13, 5
    H, mu, anh, over = synth.twodir_system(N=N, nframes=60, seed=49)
    non = mk(N, T1, T3, 8, 0, 3, H, lifetime=300.0)
    ref = orc.twodir(non, H, mu)
    non = mk(N, T1, T3, 8, 0, 3, H, lifetime=300.0)
    R = abi.twodir_batched(non, H, mu)
    assert rel(R, ref) < TOL


def test_long_t1_two_chunks_vs_oracle(orc):
    """T1 = 70: the GB/SE kernel stages the history rows of 64 t1 at a
    time, so the second chunk holds 6 rows."""
    say("N=4 T1=70 two t1 chunks")
    N, T1, T3 = 4, 70, 3
    H, mu, anh, over = synth.twodir_system(N=N, nframes=90, seed=54)
    non = mk(N, T1, T3, 5, 0, 3, H, lifetime=300.0)
    ref = orc.twodir(non, H, mu)
    non = mk(N, T1, T3, 5, 0, 3, H, lifetime=300.0)
    R = abi.twodir_batched(non, H, mu)
    assert rel(R, ref) < TOL


@pytest.mark.parametrize("tech", ["2DIR", "2DUVvis"])
def test_ring48_one_vector_per_wave_vs_oracle(orc, tech):
    """N = 48 ring: 48 kept couplings, so the two-exciton vectors (1176
    states) stay in LDS with one vector per wave (64 lanes over the sites)."""
    say(f"N=48 ring {tech}")
    N, T1, T3 = 48, 5, 6
    H, mu = synth.ring_2dir(N=N, nframes=40, seed=52)
    non = mk(N, T1, T3, 6, 0, 3, H, technique=tech, lifetime=300.0)
    ref = orc.twodir(non, H, mu, acc64=True)
    non = mk(N, T1, T3, 6, 0, 3, H, technique=tech, lifetime=300.0)
    R = abi.twodir_batched(non, H, mu)
    assert rel(R, ref) < TOL


def test_wide_system_unstaged_vs_oracle(orc):
    """N = 160: a doubles vector (12880 states) exceeds the EA kernel's LDS
    staging budget, so it reads the vectors from device memory.  The
    responses grow with N and the imaginary planes are small differences of
    them: the reference's FP32 running sums over items quantise those planes
    at about 2^-15 (the FP32 and FP64 sums of the same per-item planes differ
    by 7e-5 relative L2), so the oracle sums its per-item FP32 planes in FP64
    here."""
    say("N=160 unstaged EA")
    N, T1, T3 = 160, 3, 3
    H, mu, anh, over = synth.twodir_system(N=N, nframes=30, seed=48)
    non = mk(N, T1, T3, 9, 0, 2, H)
    ref = orc.twodir(non, H, mu, acc64=True)
    non = mk(N, T1, T3, 9, 0, 2, H)
    R = abi.twodir_batched(non, H, mu)
    assert rel(R, ref) < TOL


def test_device_records_match_host():
    import torch
    say("device records")
    N, T1, T3 = 6, 6, 6
    H, mu, anh, over = synth.twodir_system(N=N, nframes=60, seed=46)
    non = mk(N, T1, T3, 5, 0, 4, H)
    Rh = abi.twodir_batched(non, H, mu)
    Hd, mud = torch.from_numpy(H).cuda(), torch.from_numpy(mu).cuda()
    Rd = abi.twodir_batched(non, Hd, mud)
    torch.cuda.synchronize()
    assert np.array_equal(Rh, Rd)


def test_errors():
    say("error codes")
    H, mu, anh, over = synth.twodir_system(N=3, nframes=40, seed=47)
    L = abi.lib()
    R = np.zeros((12, 4, 4), np.float32)
    for kw, code in [(dict(tmax2=-1), abi.NISE_ERR_ARG), (dict(propagation=3), abi.NISE_ERR_ARG),
                     (dict(tmax2=40), abi.NISE_ERR_RANGE),
                     (dict(technique="2DIRraman"), abi.NISE_ERR_ARG),
                     (dict(end=30), abi.NISE_ERR_RANGE),
                     (dict(anharmonicity=0.0), abi.NISE_ERR_ARG)]:
        base = dict(singles=3, tmax1=4, tmax3=4, sample=2, begin=0, end=3, ts=1, propagation=1,
                    deltat=2.0, min1=1550.0, max1=1750.0, technique="2DIR", anharmonicity=16.0)
        base.update(kw)
        non = abi.make_non(**base)
        rc = L.nise_2dir_batched(C.byref(non), abi.ptr(H), H.shape[1], abi.ptr(mu), mu.shape[1],
                                 H.shape[0], None, None, None, abi.NISE_HOST, R)
        assert rc == code, (kw, rc, L.nise_last_error())


def _print2d_expected(non, Rsum, samples):
    """src/MPI_subs.c:20-34 applied to the oracle sums."""
    out = {}
    names = ["RparI.dat", "RparII.dat", "RperI.dat", "RperII.dat", "RcroI.dat", "RcroII.dat"]
    for p in range(3):
        for k in range(2):
            re = Rsum[4 * p + 2 * k].astype(np.float32) / np.float32(samples)
            im = Rsum[4 * p + 2 * k + 1].astype(np.float32) / np.float32(samples)
            re[0] /= 2
            im[0] /= 2
            rows = []
            for t1 in range(non.tmax1):
                for t3 in range(non.tmax3):
                    rows.append([t1 * non.deltat, 0.0, t3 * non.deltat, re[t3, t1], im[t3, t1]])
            out[names[2 * p + k]] = np.array(rows)
    return out


def test_calc_2dir_files(orc, tmp_path, monkeypatch):
    say("calc_2DIR files")
    L = abi.lib()
    L.calc_2DIR.argtypes = [C.POINTER(abi.TNon), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                            C.c_int]
    L.calc_2DIR.restype = None
    N, T1, T3, sr = 5, 6, 6, 5
    H, mu, anh, over = synth.twodir_system(N=N, nframes=70, seed=48)
    d = str(tmp_path)
    for name, arr in (("H.bin", H), ("mu.bin", mu), ("A.bin", anh), ("O.bin", over)):
        arr.tofile(os.path.join(d, name))
    kw = dict(length=H.shape[0], anharmonicity=0.0, lifetime=300.0,
              energyFName=os.path.join(d, "H.bin"), dipoleFName=os.path.join(d, "mu.bin"),
              anharFName=os.path.join(d, "A.bin"), overdipFName=os.path.join(d, "O.bin"))
    non = mk(N, T1, T3, sr, 2, 0, H, **kw)
    total = (H.shape[0] - T1 - T3 - 1) // sr + 1
    ref_non = mk(N, T1, T3, sr, 2, total, H, **kw)
    ref = orc.twodir(ref_non, H, mu, A_rec=anh, over_rec=over)
    monkeypatch.chdir(d)
    L.calc_2DIR(C.byref(non), 0, 1, 0, 1, 0, 0)
    exp = _print2d_expected(non, ref, total - 2)
    for name, e in exp.items():
        got = np.loadtxt(os.path.join(d, name), ndmin=2)
        assert got.shape == e.shape
        assert np.allclose(got[:, :3], e[:, :3])
        assert rel(got[:, 3:], e[:, 3:]) < 1e-4, name


def test_repeated_runs_reuse_buffers():
    """Two runs of the N = 32 ring at full size (400 samples, t1 = t3 = 64)
    after a small one: the device buffers are kept between calls, and the
    seed and history conveyors of a chunk read separate slot tables, so the
    runs agree bit for bit."""
    import torch
    say("repeated N=32 ring runs")
    N, T, sr, ns = 32, 64, 10, 400
    H, mu = synth.ring_2dir(N=N, nframes=(ns - 1) * sr + 2 * T + 2, seed=53)
    Hd, mud = torch.from_numpy(H).cuda(), torch.from_numpy(mu).cuda()
    small = abi.twodir_batched(mk(N, T, T, sr, 0, 2, H), Hd, mud)
    R1 = abi.twodir_batched(mk(N, T, T, sr, 0, ns, H), Hd, mud)
    R2 = abi.twodir_batched(mk(N, T, T, sr, 0, ns, H), Hd, mud)
    again = abi.twodir_batched(mk(N, T, T, sr, 0, 2, H), Hd, mud)
    torch.cuda.synchronize()
    assert np.array_equal(R1, R2) and np.array_equal(small, again)
