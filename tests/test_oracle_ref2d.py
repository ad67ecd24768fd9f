"""The oracle against the reference's own 2D drivers and LAPACK primitives.

tests/golden/reference_2d.npz holds the R*.dat files that the reference's
calc_2DIR / calc_2DES (compiled unmodified with MPICH and MKL, run as an MPI
singleton) wrote for small coupled systems: rings and a chain, fixed and
per-frame anharmonicities with overtone dipoles, Trotter 1 and 2, each
pathway technique, a Coupling-type Hamiltonian, 2DUVvis, a waiting time
t2 > 0 and the reference's default (sparse) propagation.

tests/golden/reference_lapack.npz holds the reference's propagate_vec_DIA,
propagate_vec_DIA_S, propagate_t2_DIA and time_evolution_mat (ssyev_ from
MKL, FP32) and propagate_double_sparce(_ES) outputs.

Tolerances: Trotter-scheme cases 2e-6 relative L2 over the six files (the
oracle's FP32 operations are the reference's; the files carry 7 significant
digits).  Cases that diagonalise (t2 > 0, sparse scheme) 1e-4: the
reference's FP32 ssyev_ eigenvectors differ from the oracle's FP64 ones by
~1e-7 per element per step, and the GB, SE and EA pathways cancel to a few
per cent of each pathway, which magnifies that difference in the summed
planes to ~2e-5 (measured).  The sparse two-exciton step on a given element
list is bit-identical.
"""
import os

import numpy as np
import pytest

import ref2d_cases as RC

GOLD_L = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_lapack.npz")
DIAG_CASES = {"ring6_t2", "ring5_2duvvis_t2", "ring6_sparse", "ring5_sparse_t2", "ring5_sparse_2duvvis"}


@pytest.mark.parametrize("name", RC.names())
def test_oracle_2d_against_reference_files(orc, name):
    non, rec, deck, R = RC.case(name)
    ref = orc.twodir(non, rec["Energy"], rec["Dipole"], C_rec=rec.get("Coupling"),
                     A_rec=rec.get("Anharmonicity"), over_rec=rec.get("OvertoneDipole"))
    exp = RC.print2d(non, ref, non.end - non.begin)
    assert exp.shape == R.shape
    assert np.allclose(exp[..., :3], R[..., :3])
    e = RC.rel(exp[..., 3:], R[..., 3:])
    assert e < (1e-4 if name in DIAG_CASES else 2e-6), f"{name}: rel L2 {e:.2e}"


def _lapack():
    return dict(np.load(GOLD_L, allow_pickle=False))


def test_dia_against_reference_lapack(orc):
    d = _lapack()
    O = orc.oracle()
    for i in range(3):
        N, sign, dt = d[f"dia{i}_args"]
        N, sign = int(N), int(sign)
        H = d[f"dia{i}_H"]
        r, q = d[f"dia{i}_in"].astype(np.float64)
        for _ in range(2):
            O.or_DIA_f64(N, H, r, q, sign, dt)
        out = d[f"dia{i}_out"]
        err = max(np.abs(r - out[0]).max(), np.abs(q - out[1]).max()) / np.abs(out).max()
        assert err < 5e-6, (i, err)
        # and not the identity or the other direction
        assert np.abs(d[f"dia{i}_in"] - out).max() > 100 * err * np.abs(out).max()


def test_dia_s_against_reference_lapack(orc):
    d = _lapack()
    O = orc.oracle()
    for i in range(3):
        N, sign, dt, thres, kept = d[f"dias{i}_args"]
        N, sign = int(N), int(sign)
        r, q = d[f"dias{i}_in"].astype(np.float64)
        k = O.or_DIA_S_f64(N, d[f"dias{i}_H"], r, q, sign, dt, np.float32(thres))
        assert k == int(kept), (i, k, kept)
        out = d[f"dias{i}_out"]
        err = max(np.abs(r - out[0]).max(), np.abs(q - out[1]).max()) / np.abs(out).max()
        assert err < 5e-6, (i, err)


def test_t2_step_against_reference_lapack(orc):
    """propagate_t2_DIA is the exact exponential applied to the t1-free
    vector and every t1 row (src/propagate.c:130-180)."""
    d = _lapack()
    O = orc.oracle()
    for i in range(2):
        N, T1, dt = d[f"t2_{i}_args"]
        N, T1 = int(N), int(T1)
        v = d[f"t2_{i}_in"].astype(np.float64)
        for H in d[f"t2_{i}_H"]:
            for row in range(T1 + 1):
                r, q = v[0, row].copy(), v[1, row].copy()
                O.or_DIA_f64(N, H, r, q, 1, dt)
                v[0, row], v[1, row] = r, q
        out = d[f"t2_{i}_out"]
        assert np.abs(v - out).max() / np.abs(out).max() < 5e-6


def test_time_evolution_mat_against_reference_lapack(orc):
    d = _lapack()
    O = orc.oracle()
    for i in range(4):
        N, dt, thres, ts, es, anh, el = d[f"sp{i}_args"]
        N, ts, el = int(N), int(ts), int(el)
        Ur, Ui = np.zeros(N * N, np.float32), np.zeros(N * N, np.float32)
        R, C = np.zeros(N * N, np.int32), np.zeros(N * N, np.int32)
        k = O.or_time_evolution_mat(N, d[f"sp{i}_H"], dt, ts, np.float32(thres), Ur, Ui, R, C)
        assert k == el
        # the caller's swap: the reference's list is (Rs, Cs) = (C, R) here
        Rs, Cs = d[f"sp{i}_RC"]
        assert np.array_equal(Rs, C[:k]) and np.array_equal(Cs, R[:k])
        U = d[f"sp{i}_U"]
        assert max(np.abs(Ur[:k] - U[0]).max(), np.abs(Ui[:k] - U[1]).max()) < 2e-6


def test_double_sparce_bit_exact(orc):
    """The sparse two-exciton step on the reference's own element list
    (propagate_double_sparce and the two-level form, two applications)."""
    d = _lapack()
    O = orc.oracle()
    for i in range(4):
        N, dt, thres, ts, es, anh, el = d[f"sp{i}_args"]
        N, ts, es, el = int(N), int(ts), int(es), int(el)
        Ur, Ui = (np.ascontiguousarray(x) for x in d[f"sp{i}_U"])
        Rs, Cs = (np.ascontiguousarray(x) for x in d[f"sp{i}_RC"])
        fr, fi = (np.ascontiguousarray(x) for x in d[f"sp{i}_in"].copy())
        Anh = np.ascontiguousarray(d[f"sp{i}_Anh"])
        for _ in range(2):
            O.or_double_sparce_f32(N, Ur, Ui, Rs, Cs, fr, fi, el, ts, Anh.ctypes.data, anh, dt, es)
        out = d[f"sp{i}_out"]
        assert np.array_equal(fr, out[0]) and np.array_equal(fi, out[1]), i
