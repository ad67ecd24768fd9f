"""GPU parity on the benchmark workloads themselves, each against the oracle.

Every workload below is the one bench.py times, built by the same generator
with the same seeds and cuts; where the full run would take the CPU oracle
too long, a subset of its start samples is run (all propagation steps, all
polarisations, the same plan code):

  N = 16 chain absorption          all 200 samples (tmax 128)
  N = 256 16x16 dipole lattice     32 samples x 3 pols x 256 steps, rows 16
                                   (the bench's conveyor shape) and the
                                   default shape
  N = 128 chain adjoint            6 samples: gradient against central
                                   finite differences of the FP64 oracle
                                   for K = 16, 32, 64
  96-system uneven batch           every system, 2 samples x 3 pols x 512
                                   steps, N up to 981 (Trotter) and N <= 64
                                   (exact exponential)
  N = 32 ring 2DIR                 t1 = t3 = 64, 2 samples x 21 items
  2DIR over more than one chunk    1100 samples (the batch splits into
                                   chunks of at most 1024 samples)

Tolerances: absorption 1e-5 relative L2 on the response (the Trotter
propagation follows the reference's FP32 operation order; sums are FP64
here, FP32 running sums in the oracle), exact-exponential systems 1e-5
(FP64 matrix exponential on the GPU, FP64 eigensystem in the oracle, FP32
vectors on both sides); adjoint 1e-3 relative
against finite differences; 2DIR 1e-4 relative L2 over the 12 planes.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from nise_2017_amd import abi, adjoint, synth

pytestmark = pytest.mark.gpu

SEED0 = synth.SEED0
THREADS = max(1, min(16, os.cpu_count() or 1))


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    dt = np.complex128 if np.iscomplexobj(a) or np.iscomplexobj(b) else np.float64
    a, b = a.astype(dt), b.astype(dt)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def say(msg):
    print(f"[configs] {msg}", flush=True)


def centred(non, lo, hi):
    non.shifte = np.float32((np.float32(hi) + np.float32(lo)) / np.float32(2))
    return non


def test_chain16_all_samples(orc):
    """N = 16 chain, Full type, 4200 frames, tmax 128, 200 samples every 20
    frames, Trotter 1: the whole run."""
    N, T, sr, ns = 16, 128, 20, 200
    H, mu = synth.chain_absorption(N=N, nframes=4200, seed=SEED0)
    lo, hi = synth.shifted_center(H, N)
    mk = lambda: centred(abi.make_non(singles=N, tmax1=T, sample=sr, begin=0, end=ns, ts=1,  # noqa: E731
                                      propagation=1, deltat=2.0, min1=lo, max1=hi, fft=2048,
                                      lifetime=1300.0), lo, hi)
    say("N=16 chain, 200 samples")
    p = abi.AbsPlan(mk(), H, H.shape[1], mu, mu.shape[1], H.shape[0])
    re, im = p.execute()
    p.close()
    o_re, o_im = orc.absorption(mk(), H, mu, nthreads=THREADS)
    e = rel(re + 1j * im, o_re + 1j * o_im)
    say(f"rel L2 {e:.2e}")
    assert e < 1e-5


@pytest.mark.parametrize("rows", [16, 0])
def test_lattice256_subset(orc, rows):
    """N = 256 lattice of the bench (16 x 16, J = 8 (a/r)^3 kappa cut at 0.1
    cm^-1, AR(1) energies, seed 20261202 + 1), Full FP32 frames: the first
    32 start samples (every 10 frames) x 3 pols x 256 steps."""
    import torch
    L, T, sr, ns = 16, 256, 10, 32
    N = L * L
    nfr = (ns - 1) * sr + T + 2
    rng, mus, coup = synth.lattice_static(L, seed=SEED0 + 1)
    diag = synth.lattice_diag(rng, nfr, N)
    H = synth.full_records(diag, coup)
    mu = synth.dipole_records(mus, nfr)
    lo, hi = float(diag.mean()) - 100.0, float(diag.mean()) + 100.0
    mk = lambda: centred(abi.make_non(singles=N, tmax1=T, sample=sr, begin=0, end=ns, ts=1,  # noqa: E731
                                      propagation=1, deltat=2.0, min1=lo, max1=hi, couplingcut=0.0,
                                      fft=2048), lo, hi)
    say(f"N=256 lattice, {ns} samples, rows {rows or 'default'}")
    Hd, mud = torch.from_numpy(H).cuda(), torch.from_numpy(mu).cuda()
    p = abi.AbsPlan(mk(), Hd, H.shape[1], mud, mu.shape[1], nfr, where=abi.NISE_DEVICE)
    if rows:
        p.set_rows(rows)
    info = p.info()
    re, im = p.execute(flags=2)  # the bench replays the captured graph
    re2, im2 = p.execute(flags=2)
    p.close()
    assert np.array_equal(re, re2) and np.array_equal(im, im2)
    o_re, o_im = orc.absorption(mk(), H, mu, nthreads=THREADS)
    e = rel(re + 1j * im, o_re + 1j * o_im)
    say(f"plan {info}: rel L2 {e:.2e}")
    assert e < 1e-5


@pytest.mark.parametrize("K", [16, 32, 64])
def test_chain128_adjoint_fd(orc, K):
    """N = 128 chain of the adjoint workload (20,000 frames generated, seed
    20261202 + 3, tmax 256, samples every 19 frames, lifetime 1300 fs, FFT
    2048), 6 samples: dL/d(offset) for the three offsets with the largest
    gradient and dL/d(scale), against central differences of the FP64
    oracle (h = 0.01 cm^-1, 1e-4)."""
    N, T, sr, ns, nfr = 128, 256, 19, 6, 20000
    H, mu = synth.chain_absorption(N=N, nframes=nfr, seed=SEED0 + 3)
    lo, hi = synth.shifted_center(H[:(1000 - 1) * sr + T], N)  # the full run's window
    need = (ns - 1) * sr + T + 1
    H, mu = np.ascontiguousarray(H[:need]), np.ascontiguousarray(mu[:need])
    non = centred(abi.make_non(singles=N, tmax1=T, sample=sr, begin=0, end=ns, ts=1, propagation=1,
                               deltat=2.0, min1=lo, max1=hi, fft=2048, lifetime=1300.0), lo, hi)
    freq, A_re, A_im = adjoint.spectrum_matrix(non, ns)
    rng = np.random.default_rng(SEED0 + 3)
    d_t = rng.normal(0, 5, N)
    St = orc.absorption_param_f64(non, H, mu, d_t, 1.1, nthreads=THREADS)
    target = A_re @ St[0] + A_im @ St[1]
    say(f"N=128 chain adjoint, K={K}")
    model = adjoint.AdjointAbsorption(non, H, mu, every=K)
    S = model.forward(None, 1.0)
    r = A_re @ S[0] + A_im @ S[1] - target
    gd, gc = model.backward(2 * A_re.T @ r, 2 * A_im.T @ r)
    model.close()

    def loss(d, c):
        re, im = orc.absorption_param_f64(non, H, mu, d, c, nthreads=THREADS)
        rr = A_re @ re + A_im @ im - target
        return float(rr @ rr)
    idx = list(np.argsort(-np.abs(gd))[:3])
    d0 = np.zeros(N)
    fd = []
    for a in idx:
        e = np.zeros(N)
        e[a] = 0.01
        fd.append((loss(d0 + e, 1.0) - loss(d0 - e, 1.0)) / 0.02)
    fd.append((loss(d0, 1.0 + 1e-4) - loss(d0, 1.0 - 1e-4)) / 2e-4)
    g = np.array([gd[a] for a in idx] + [gc])
    err = rel(g, np.array(fd))
    say(f"K={K}: gradient {g}, finite differences {np.array(fd)}, rel {err:.2e}")
    assert err < 1e-3


def _uneven_systems(samples):
    specs = synth.uneven_batch(samples=samples)
    systems = []
    for s in specs:
        non = centred(abi.make_non(singles=s["N"], tmax1=512, sample=10, begin=0, end=samples, ts=1,
                                   propagation=s["propagation"], deltat=1.0, min1=s["min1"],
                                   max1=s["max1"], hamiltonian="Coupling", couplingcut=0.5, thres=0.0),
                      s["min1"], s["max1"])
        systems.append((non, s["H"], s["mu"], s["C"]))
    return systems


def oracle_many(orc, systems):
    """The oracle's absorption of every system, longest jobs first on a
    thread pool (the oracle releases the interpreter lock)."""
    def cost(s):
        N = s[0].singles
        return N ** 3 if s[0].propagation == 0 else N * N
    order = sorted(range(len(systems)), key=lambda k: -cost(systems[k]))
    out = [None] * len(systems)

    def run(k):
        non, H, mu, Cr = systems[k]
        out[k] = orc.absorption(non, H, mu, C_rec=Cr)
    with ThreadPoolExecutor(THREADS) as ex:
        list(ex.map(run, order))
    return out


def test_uneven_batch_all_systems(orc):
    """All 96 systems of the uneven batch in one call, 2 samples each."""
    import torch
    systems = _uneven_systems(2)
    Ns = [s[0].singles for s in systems]
    say(f"96-system batch, sum N {sum(Ns)}, max N {max(Ns)}")
    dev = [(non, torch.from_numpy(H).cuda(), torch.from_numpy(mu).cuda(), torch.from_numpy(Cr).cuda())
           for non, H, mu, Cr in systems]
    out, done = abi.absorption_multi(dev)
    ref = oracle_many(orc, systems)
    worst = 0.0
    for (non, _, _, _), (re, im), (o_re, o_im) in zip(systems, out, ref):
        e = rel(re + 1j * im, o_re + 1j * o_im)
        worst = max(worst, e)
        assert e < 1e-5, f"N={non.singles} propagation {non.propagation}: rel L2 {e:.2e}"
    say(f"worst rel L2 {worst:.2e}")
    assert np.all(done > 0)


def test_ring32_2dir_two_samples(orc):
    """The 2DIR workload (N = 32 ring, seed 20261202 + 2, anharmonicity 16,
    t1 = t3 = 64, t2 = 0, samples every 10 frames, Trotter 1) on its first
    2 samples x 21 items."""
    N, T, sr, ns = 32, 64, 10, 2
    H, mu = synth.ring_2dir(N=N, nframes=(400 - 1) * sr + 2 * T + 2, seed=SEED0 + 2)
    lo, hi = synth.shifted_center(H, N)  # the full run's window
    need = (ns - 1) * sr + 2 * T + 2
    H, mu = np.ascontiguousarray(H[:need]), np.ascontiguousarray(mu[:need])
    mk = lambda: abi.make_non(singles=N, tmax1=T, tmax2=0, tmax3=T, sample=sr, begin=0, end=ns,  # noqa: E731
                              ts=1, propagation=1, deltat=2.0, min1=lo, max1=hi, technique="2DIR",
                              anharmonicity=16.0)
    say("N=32 ring 2DIR, 2 samples")
    R = abi.twodir_batched(mk(), H, mu)
    ref = orc.twodir(mk(), H, mu, acc64=True)
    e = rel(R, ref)
    say(f"rel L2 {e:.2e}")
    assert e < 1e-4


def test_2dir_more_than_one_chunk(orc):
    """1100 samples: the batch is cut into chunks of at most 1024 samples
    (kept buffers, per-chunk slot tables, accumulation across chunks)."""
    N, T, sr, ns = 5, 6, 2, 1100
    H, mu, anh, over = synth.twodir_system(N=N, nframes=(ns - 1) * sr + 2 * T + 3, seed=71)
    lo, hi = synth.shifted_center(H, N)
    mk = lambda: abi.make_non(singles=N, tmax1=T, tmax2=0, tmax3=T, sample=sr, begin=0, end=ns,  # noqa: E731
                              ts=1, propagation=1, deltat=2.0, min1=lo, max1=hi, technique="2DIR",
                              anharmonicity=16.0, lifetime=300.0)
    say(f"2DIR N={N}, {ns} samples")
    R = abi.twodir_batched(mk(), H, mu)
    ref = orc.twodir(mk(), H, mu, acc64=True)
    e = rel(R, ref)
    say(f"rel L2 {e:.2e}")
    assert e < 1e-4
