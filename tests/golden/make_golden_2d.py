"""Write tests/golden/reference_2d.npz from the reference's own 2D drivers,
and tests/golden/reference_lapack.npz from its exact-exponential and sparse
primitives.

Runs only where the reference tree and MPICH/MKL are present (this build
container): oracle/_ref/nise_ref2d is the reference's calc_2DIR.c /
calc_2DES.c with the rest of its sources, compiled unmodified by
oracle/Makefile (target ref2d) and run here as an MPI singleton.  For every
case this script writes the trajectory files and an input deck into a
temporary directory, runs the program there and keeps the six R*.dat files
it writes.  The fixtures are data: the input records, the deck, and the
response planes as print2D wrote them (divided by the sample count, t3 = 0
row halved, %e text).

The primitives come from oracle/_ref/libnise_ref_lapack.so: the
reference's propagate.c with diagonalizeLPD calling MKL's ssyev_
(propagate_vec_DIA, propagate_vec_DIA_S, propagate_t2_DIA,
time_evolution_mat, propagate_double_sparce and its two-level form).

    make -C oracle ref && python tests/golden/make_golden_2d.py
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from nise_2017_amd import synth  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_2d.npz")
OUT_LAPACK = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_lapack.npz")
LAPACK_SO = os.path.join(ROOT, "oracle", "_ref", "libnise_ref_lapack.so")
ICM2IFS = np.float32(2.99792458e-5)
TWOPI = np.float32(2 * 3.14159265359)
PROG = os.path.join(ROOT, "oracle", "_ref", "nise_ref2d")
FILES = ("RparI.dat", "RparII.dat", "RperI.dat", "RperII.dat", "RcroI.dat", "RcroII.dat")


def ring_or_chain(N, nframes, seed, ring, jmean=-6.0, jsd=2.0, sd=12.0, dip_fluct=0.15, dt=2.0):
    """Full records for N sites (AR(1) energies 1650 / sd, static nearest-
    neighbour couplings N(jmean, jsd) on a chain or ring plus AR(1)
    fluctuations of sd 1), fluctuating dipoles, per-frame anharmonicities
    and overtone dipoles."""
    H, mu, anh, over = synth.twodir_system(N=N, nframes=nframes, seed=seed, sd=sd, coupled=True,
                                           jmean=jmean, jsd=jsd, dip_fluct=dip_fluct, dt=dt)
    if ring and N > 2:
        rng = np.random.default_rng(seed + 1)
        a, b = np.triu_indices(N)
        k = 1 + np.nonzero((a == 0) & (b == N - 1))[0][0]
        H[:, k] = np.float32(rng.normal(jmean, jsd))
    return H, mu, anh, over


def deck_text(d):
    lines = []
    for k, v in d.items():
        if isinstance(v, (list, tuple)):
            v = " ".join(str(x) for x in v)
        lines.append(f"{k} {v}")
    return "\n".join(lines) + "\n"


def run_case(name, deck, files):
    with tempfile.TemporaryDirectory() as d:
        for fname, arr in files.items():
            arr.astype(np.float32).tofile(os.path.join(d, fname))
        with open(os.path.join(d, "input2D"), "w") as fh:
            fh.write(deck_text(deck))
        env = dict(os.environ, OMP_NUM_THREADS="1")
        p = subprocess.run([PROG, "input2D"], cwd=d, capture_output=True, text=True, env=env, timeout=600)
        if p.returncode != 0:
            raise RuntimeError(f"{name}: reference run failed ({p.returncode}):\n{p.stdout[-3000:]}\n{p.stderr}")
        planes = []
        for f in FILES:
            t = np.loadtxt(os.path.join(d, f), ndmin=2)
            planes.append(t)
        return np.stack(planes), p.stdout


def ring_frame(rng, N, esd=12.0, jmean=-6.0, jsd=2.0, extra=0.0):
    """Packed upper triangle: site energies N(0, esd), ring couplings
    N(jmean, jsd), plus N(0, extra) on every other pair when extra > 0."""
    H = np.zeros(N * (N + 1) // 2, np.float32)
    for a in range(N):
        H[synth.sindex(a, a, N)] = rng.normal(0, esd)
        for b in range(a + 1, N):
            if extra > 0:
                H[synth.sindex(a, b, N)] = rng.normal(0, extra)
        if N > 1:
            H[synth.sindex(a, (a + 1) % N, N)] = rng.normal(jmean, jsd)
    return H


def scaled_threshold(t, dt, ts):
    """readInput's rescaling of Threshold (src/readinput.c:262-264), FP32."""
    g = np.float32(np.float32(np.float32(np.float32(dt) * ICM2IFS) * TWOPI) / np.float32(ts))
    return np.float32(np.float32(t) * np.float32(g * g))


def lapack_primitives():
    import ctypes as C
    f32 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    i32 = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    L = C.CDLL(LAPACK_SO)
    L.ref_DIA.argtypes = [C.c_int, f32, f32, f32, C.c_int, C.c_float]
    L.ref_DIA_S.argtypes = [C.c_int, f32, f32, f32, C.c_int, C.c_float, C.c_float]
    L.ref_DIA_S.restype = C.c_int
    L.ref_t2_DIA.argtypes = [C.c_int, f32, f32, f32, f32, f32, C.c_int, C.c_float]
    L.ref_time_evolution_mat.argtypes = [C.c_int, f32, C.c_float, C.c_int, C.c_float, f32, f32, i32, i32]
    L.ref_time_evolution_mat.restype = C.c_int
    L.ref_double_sparce.argtypes = [C.c_int, f32, f32, i32, i32, f32, f32, C.c_int, C.c_int, f32, C.c_float,
                                    C.c_float, C.c_int]
    rng = np.random.default_rng(20261203)
    out = {}
    # propagate_vec_DIA: (N, sign, dt, extra couplings), two steps
    for i, (N, sign, dt, extra) in enumerate([(3, 1, 2.0, 0.0), (6, -1, 2.0, 0.0), (17, 1, 5.0, 3.0)]):
        H = ring_frame(rng, N, extra=extra)
        cr, ci = rng.normal(size=N).astype(np.float32), rng.normal(size=N).astype(np.float32)
        r, q = cr.copy(), ci.copy()
        for _ in range(2):
            L.ref_DIA(N, H, r, q, sign, dt)
        out[f"dia{i}_args"] = np.array([N, sign, dt], np.float64)
        out[f"dia{i}_H"], out[f"dia{i}_in"], out[f"dia{i}_out"] = H, np.stack([cr, ci]), np.stack([r, q])
    # propagate_vec_DIA_S: (N, sign, dt, Threshold keyword, Trotter), one step
    for i, (N, sign, dt, t, ts) in enumerate([(6, 1, 2.0, 0.001, 1), (12, -1, 10.0, 0.01, 1), (8, 1, 4.0, 0.1, 2)]):
        H = ring_frame(rng, N, extra=1.0)
        thres = scaled_threshold(t, dt, ts)
        cr, ci = rng.normal(size=N).astype(np.float32), rng.normal(size=N).astype(np.float32)
        r, q = cr.copy(), ci.copy()
        kept = L.ref_DIA_S(N, H, r, q, sign, dt, thres)
        out[f"dias{i}_args"] = np.array([N, sign, dt, thres, kept], np.float64)
        out[f"dias{i}_H"], out[f"dias{i}_in"], out[f"dias{i}_out"] = H, np.stack([cr, ci]), np.stack([r, q])
    # propagate_t2_DIA: one vector and T1 rows, three waiting-time steps
    for i, (N, T1, dt) in enumerate([(6, 3, 2.0), (9, 2, 5.0)]):
        Hs = [ring_frame(rng, N) for _ in range(3)]
        cr, ci = rng.normal(size=N).astype(np.float32), rng.normal(size=N).astype(np.float32)
        vr = rng.normal(size=(T1, N)).astype(np.float32)
        vi = rng.normal(size=(T1, N)).astype(np.float32)
        r, q, a, b = cr.copy(), ci.copy(), vr.copy(), vi.copy()
        for H in Hs:
            L.ref_t2_DIA(N, H, r, q, a, b, T1, dt)
        out[f"t2_{i}_args"] = np.array([N, T1, dt], np.float64)
        out[f"t2_{i}_H"] = np.stack(Hs)
        out[f"t2_{i}_in"] = np.concatenate([np.stack([cr, ci])[:, None, :], np.stack([vr, vi])], axis=1)
        out[f"t2_{i}_out"] = np.concatenate([np.stack([r, q])[:, None, :], np.stack([a, b])], axis=1)
    # time_evolution_mat and the sparse doubles step on its element list, as
    # propagate_double_2DIR_control calls them (row and column arrays
    # swapped, src/propagate.c:1496); (N, dt, Threshold, Trotter, es, fixed
    # anharmonicity or 0 for per-site values)
    for i, (N, dt, t, ts, es, anh) in enumerate([(4, 2.0, 0.001, 1, 0, 16.0), (6, 2.0, 0.001, 2, 0, 0.0),
                                                   (5, 10.0, 0.01, 1, 1, 0.0), (7, 5.0, 0.0, 1, 0, 20.0)]):
        H = ring_frame(rng, N, extra=0.5)
        thres = scaled_threshold(t, dt, ts) if t > 0 else np.float32(0)
        Ur, Ui = np.zeros(N * N, np.float32), np.zeros(N * N, np.float32)
        Rs, Cs = np.zeros(N * N, np.int32), np.zeros(N * N, np.int32)
        el = L.ref_time_evolution_mat(N, H, dt, ts, thres, Ur, Ui, Cs, Rs)
        nn = N * (N + 1) // 2
        fr, fi = rng.normal(size=nn).astype(np.float32), rng.normal(size=nn).astype(np.float32)
        Anh = rng.normal(16, 3, N).astype(np.float32)
        a, b = fr.copy(), fi.copy()
        for _ in range(2):
            L.ref_double_sparce(N, Ur, Ui, Rs, Cs, a, b, el, ts, Anh, anh, dt, es)
        out[f"sp{i}_args"] = np.array([N, dt, thres, ts, es, anh, el], np.float64)
        out[f"sp{i}_H"], out[f"sp{i}_Anh"] = H, Anh
        out[f"sp{i}_U"] = np.stack([Ur[:el], Ui[:el]])
        out[f"sp{i}_RC"] = np.stack([Rs[:el], Cs[:el]])
        out[f"sp{i}_in"], out[f"sp{i}_out"] = np.stack([fr, fi]), np.stack([a, b])
    np.savez_compressed(OUT_LAPACK, **out)
    print(f"wrote {OUT_LAPACK} ({os.path.getsize(OUT_LAPACK)} bytes)")


def main():
    lapack_primitives()
    if not os.path.exists(PROG):
        raise SystemExit(f"{PROG} missing: make -C oracle ref2d")
    out = {}
    cases = []
    base = {"Propagation": "Coupling", "Couplingcut": 0, "Hamiltonianfile": "Energy.bin",
            "Dipolefile": "Dipole.bin", "Timestep": 2, "Trotter": 1, "Anharmonicity": 16,
            "Lifetime": 300, "Technique": "2DIR", "FFT": 64, "BeginPoint": 0}

    def add(name, N, T1, T2, sr, ns, seed, ring=True, deck_over=None, anh_files=False,
            coupling_type=False, nframes=None):
        nfr = nframes or (ns - 1) * sr + T1 + T2 + T1 + 4
        H, mu, anh, over = ring_or_chain(N, nfr, seed, ring)
        lo, hi = synth.shifted_center(H, N)
        deck = dict(base)
        deck.update({"Length": nfr, "Samplerate": sr, "EndPoint": ns, "RunTimes": (T1, T2, T1),
                     "Singles": N, "MinFrequencies": (round(lo, 3),) * 3, "MaxFrequencies": (round(hi, 3),) * 3})
        files = {"Energy.bin": H, "Dipole.bin": mu}
        if anh_files:
            deck["Anharmonicity"] = 0
            deck["Anharmonicfile"] = "Anharmonicity.bin"
            deck["Overtonedipolefile"] = "OvertoneDipole.bin"
            files["Anharmonicity.bin"], files["OvertoneDipole.bin"] = anh, over
        if coupling_type:
            a, b = np.triu_indices(N)
            Hc = np.ascontiguousarray(np.concatenate([H[:, :1], H[:, 1:][:, a == b]], axis=1))
            Cr = np.ascontiguousarray(H[0])
            deck["HamiltonianType"] = "Coupling"
            deck["Couplingfile"] = "Coupling.bin"
            files["Energy.bin"], files["Coupling.bin"] = Hc, Cr
            files["Dipole.bin"] = np.ascontiguousarray(mu[:1])
            H, mu = Hc, files["Dipole.bin"]
        if deck_over:
            deck.update(deck_over)
        R, log = run_case(name, deck, files)
        k = len(cases)
        cases.append(name)
        out[f"c{k}_deck"] = np.array(json.dumps(deck))
        for fname, arr in files.items():
            out[f"c{k}_{fname.split('.')[0]}"] = np.ascontiguousarray(arr, np.float32)
        out[f"c{k}_R"] = R.astype(np.float64)
        print(f"{name}: N={N} t1={T1} t2={T2} samples={ns}, |R| max {np.abs(R[..., 3:]).max():.3e}")

    # coupled 2DIR, the library's propagation (Trotter, fixed anharmonicity)
    add("ring6_coupling", 6, 8, 0, 5, 4, 101)
    # per-frame anharmonicities and overtone dipoles, Trotter 2, chain
    add("chain5_anhfile", 5, 7, 0, 4, 3, 102, ring=False, anh_files=True, deck_over={"Trotter": 2})
    # the individual pathway techniques
    for tech in ("GBIR", "SEIR", "EAIR", "noEAIR"):
        add(f"ring4_{tech}", 4, 6, 0, 3, 3, 103, deck_over={"Technique": tech})
    # Coupling-type Hamiltonian (static couplings file), no lifetime
    add("ring5_ctype", 5, 6, 0, 4, 3, 104, coupling_type=True, deck_over={"Lifetime": 0})
    # 2DUVvis (two-level sites)
    add("ring6_2duvvis", 6, 8, 0, 5, 3, 105, deck_over={"Technique": "2DUVvis"})
    add("ring4_noEAUVvis", 4, 6, 0, 3, 3, 106, deck_over={"Technique": "noEAUVvis"})
    # waiting time t2 > 0 (propagate_t2_DIA during t2)
    add("ring6_t2", 6, 8, 5, 5, 3, 107)
    add("ring5_2duvvis_t2", 5, 6, 4, 4, 3, 108, deck_over={"Technique": "2DUVvis"})
    # the reference's default propagation (the sparse scheme; Threshold is
    # rescaled by (dt/hbar)^2 in readInput)
    add("ring6_sparse", 6, 8, 0, 5, 3, 109, deck_over={"Propagation": "Sparse", "Threshold": 0.001})
    add("ring5_sparse_t2", 5, 6, 3, 4, 3, 110, deck_over={"Propagation": "Sparse", "Threshold": 0.01})
    add("ring5_sparse_2duvvis", 5, 6, 0, 4, 3, 111,
        deck_over={"Propagation": "Sparse", "Threshold": 0.001, "Technique": "2DUVvis"})
    out["cases"] = np.array(json.dumps(cases))
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes, {len(cases)} cases)")


if __name__ == "__main__":
    main()
