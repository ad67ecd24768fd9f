"""End-to-end parity against the reference's own compiled code.

oracle/_ref/libnise_ref.so is built in the build container from the
reference's unmodified sources (propagate.c, NISE_subs.c, read_trajectory.c,
...; oracle/Makefile) and travels with the tree as a built library.  Its
absorption loop reads the same GROBIN files with the reference's read_He /
read_mue and propagates with the reference's propagate_vec_coupling_S.
The propagation is the same FP32 arithmetic (bit-exact per step, pinned by
the golden-vector tests); the response differs only in its summation: the
reference adds every product mu_i * psi_i straight into one FP32 running sum
per time step (calc_S1, src/absorption.c:205-212; 9,600 additions per step at
cfg0), this path sums in FP64.  On cfg0 that FP32 running sum alone differs
from an FP64 sum of the identical FP32 propagation by 3.4e-6 relative L2, so
the response must match to 1e-5 and the spectrum to the 1e-4 spectral
tolerance.  Skipped where the reference build is absent."""
import os

import numpy as np
import pytest

from nise_2017_amd import abi, adjoint, synth

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def ref():
    from oracle import oracle
    if not oracle.have_ref():
        pytest.skip("reference build (oracle/_ref) not present")
    return oracle.ref()


@pytest.mark.parametrize("case", ["cfg0_chain16", "lattice36_shift", "coupling_type"])
def test_absorption_vs_reference_build(ref, tmp_path, case):
    d = str(tmp_path)
    shift = None
    Crec = None
    if case == "cfg0_chain16":  # N=16 chain, tmax 128, 200 samples every 20 frames
        N, T, sr, ns = 16, 128, 20, 200
        H, mu = synth.chain_absorption(N=N, nframes=4200, seed=20261202)
        htype = b"Full"
        lo, hi = synth.shifted_center(H, N)
    else:
        N, T, sr, ns = 36, 64, 9, 30
        rng, mus, coup = synth.lattice_static(6, seed=33, cut=0.1)
        nfr = (ns - 1) * sr + T + 5
        H = synth.full_records(synth.lattice_diag(rng, nfr, N), coup)
        mu = synth.dipole_records(mus, nfr)
        htype = b"Full"
        lo, hi = synth.shifted_center(H, N)
        if case == "lattice36_shift":
            shift = [(3, 1.0, 4.0), (20, 0.95, 80.0)]
        else:
            Crec = np.ascontiguousarray(H[0])
            a, b = np.triu_indices(N)
            H = np.ascontiguousarray(np.concatenate([H[:, :1], H[:, 1:][:, a == b]], axis=1))
            mu = np.ascontiguousarray(mu[:1])
            htype = b"Coupling"
    hf, df, cf, sf = (os.path.join(d, n) for n in ("H.bin", "mu.bin", "C.bin", "shift.txt"))
    H.tofile(hf)
    mu.tofile(df)
    if Crec is not None:
        Crec.tofile(cf)
    if shift:
        with open(sf, "w") as fh:
            fh.write(f"{len(shift)}\n" + "".join(f"{s} {a} {b}\n" for s, a, b in shift))
    re_r = np.zeros(T, np.float32)
    im_r = np.zeros(T, np.float32)
    ref.ref_absorption_files(N, hf.encode(), df.encode(), cf.encode() if Crec is not None else b"",
                             sf.encode() if shift else b"", htype, lo, hi, 2.0, 0.0, 1, T, sr, 0, ns,
                             H.shape[0], re_r, im_r)
    kw = dict(singles=N, tmax1=T, sample=sr, begin=0, end=ns, ts=1, propagation=1, deltat=2.0,
              min1=lo, max1=hi, fft=2048, lifetime=1300.0)
    if shift:
        kw["single_shift"] = shift
    if Crec is not None:
        kw["hamiltonian"] = "Coupling"
    non = abi.make_non(**kw)
    non.shifte = np.float32((np.float32(hi) + np.float32(lo)) / np.float32(2))
    p = abi.AbsPlan(non, H, H.shape[1], mu, mu.shape[1], H.shape[0], Crec)
    re, im = p.execute()
    p.close()
    e_td = rel(re + 1j * im, re_r + 1j * im_r)
    _, A_re, A_im = adjoint.spectrum_matrix(non, ns)
    s_gpu = A_re @ re.astype(np.float64) + A_im @ im.astype(np.float64)
    s_ref = A_re @ re_r.astype(np.float64) + A_im @ im_r.astype(np.float64)
    e_sp = rel(s_gpu, s_ref)
    print(f"[reference] {case}: response rel L2 {e_td:.1e}, spectrum rel L2 {e_sp:.1e}", flush=True)
    assert e_td < 1e-5 and e_sp < 1e-4
