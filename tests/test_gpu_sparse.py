"""GPU parity of the sparse DIA scheme (src/propagate.c:185-258, selected by
propagation 0 with a threshold in (0, 1], :25-28) through the C ABI: the
reference-compatible propagate_vector / propagate_matrix and the batched
absorption, against the FP64 oracle on the same inputs.

The GPU builds U from its FP64 eigensystem with the phases e*f rounded to
float (as the reference does) and tests FP32 elements against the threshold;
the oracle uses exact FP64 phases.  Thresholds are therefore put inside a gap
of the |U_ab|^2 distribution of the inputs, where no element can change side,
and the tolerance is FP32's (1e-5 relative); each case also checks that the
scheme drops something visible (the sparse oracle differs from plain DIA by
more than the tolerance)."""
import ctypes as C

import numpy as np
import pytest

from nise_2017_amd import abi, synth

pytestmark = pytest.mark.gpu

F = float(np.float32(np.float32(np.float32(2.0) * np.float32(2.99792458e-5)) * np.float32(2 * 3.14159265359)))


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def dense(tri, N):
    a, b = np.triu_indices(N)
    Hd = np.zeros((N, N))
    Hd[a, b] = tri
    Hd[b, a] = tri
    return Hd


def u_sq(Hd):
    w, V = np.linalg.eigh(Hd)
    U = (V * np.exp(-1j * w * F)) @ V.T
    return (np.abs(U.astype(np.complex64)) ** 2).ravel()


def gap_threshold(values, lo, hi):
    """Geometric centre of the widest gap (in log) among values in [lo, hi]."""
    v = np.sort(values[(values > lo) & (values < hi)])
    g = np.diff(np.log(v))
    k = int(np.argmax(g))
    assert np.exp(g[k]) > 1.2, "no usable gap"
    return float(np.sqrt(v[k] * v[k + 1]))


@pytest.mark.parametrize("N", [6, 23])
def test_sparse_vector_and_matrix_vs_oracle(orc, nise_lib, N):
    rng = np.random.default_rng(100 + N)
    Hd = np.diag(rng.normal(0, 60, N))
    for a in range(N - 1):  # chain with a weak second neighbour
        Hd[a, a + 1] = Hd[a + 1, a] = rng.normal(-8, 2)
        if a + 2 < N:
            Hd[a, a + 2] = Hd[a + 2, a] = rng.normal(0, 0.5)
    a, b = np.triu_indices(N)
    tri = np.ascontiguousarray(Hd[a, b], np.float32)
    thres = gap_threshold(u_sq(dense(tri, N)), 1e-10, 1e-4)
    non = abi.make_non(singles=N, deltat=2.0, propagation=0, thres=thres, ts=1)
    v = (rng.normal(size=N) + 1j * rng.normal(size=N)).astype(np.complex64)
    g_r, g_i = np.ascontiguousarray(v.real), np.ascontiguousarray(v.imag)
    nise_lib.propagate_vector(C.byref(non), tri, g_r, g_i, 1, 1, 1)  # samples != begin: no report
    d_r, d_i = v.real.astype(np.float64), v.imag.astype(np.float64)
    kept = orc.oracle().or_DIA_S_f64(N, tri, d_r, d_i, 1, 2.0, thres)
    assert 0 < kept < N * N
    o = d_r + 1j * d_i
    assert rel(g_r + 1j * g_i, o) < 1e-5
    p_r, p_i = v.real.astype(np.float64), v.imag.astype(np.float64)
    orc.oracle().or_DIA_f64(N, tri, p_r, p_i, 1, 2.0)
    assert rel(p_r + 1j * p_i, o) > 1e-5  # the dropped elements matter
    # propagate_matrix: the N columns of an N x N block
    M = (rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))).astype(np.complex64)
    m_r, m_i = np.ascontiguousarray(M.real), np.ascontiguousarray(M.imag)
    nise_lib.propagate_matrix(C.byref(non), tri, m_r.ravel(), m_i.ravel(), 1, 1, 1)
    for j in range(N):
        c_r, c_i = M[j].real.astype(np.float64), M[j].imag.astype(np.float64)
        orc.oracle().or_DIA_S_f64(N, tri, c_r, c_i, 1, 2.0, thres)
        assert rel(m_r[j] + 1j * m_i[j], c_r + 1j * c_i) < 1e-5


def test_sparse_absorption_through_batched(orc):
    N, T, sr, ns = 16, 128, 20, 30
    H, mu = synth.chain_absorption(N=N, nframes=(ns - 1) * sr + T + 2, seed=77)
    lo, hi = synth.shifted_center(H, N)
    shift = np.float32((np.float32(hi) + np.float32(lo)) / np.float32(2))
    vals = []
    for fr in range((ns - 1) * sr + T):
        Hd = dense(H[fr, 1:].astype(np.float64), N)
        Hd[np.arange(N), np.arange(N)] -= float(shift)
        vals.append(u_sq(Hd))
    thres = gap_threshold(np.concatenate(vals), 1e-12, 1e-6)

    def make(th):
        non = abi.make_non(singles=N, tmax1=T, sample=sr, begin=0, end=ns, ts=1, propagation=0, deltat=2.0,
                           min1=lo, max1=hi, thres=th)
        non.shifte = shift
        return non

    non = make(thres)
    re = np.zeros(T, np.float32)
    im = np.zeros(T, np.float32)
    abi.check(abi.lib().nise_absorption_batched(C.byref(non), abi.ptr(H), H.shape[1], abi.ptr(mu), mu.shape[1],
                                                H.shape[0], None, abi.NISE_HOST, 0, re, im))
    o_re, o_im = orc.absorption(non, H, mu, f64=True)
    d_re, d_im = orc.absorption(make(0.0), H, mu, f64=True)
    e_gpu = rel(re + 1j * im, o_re + 1j * o_im)
    e_dia = rel(o_re + 1j * o_im, d_re + 1j * d_im)
    print(f"[sparse] thres {thres:.3e}: GPU vs sparse oracle {e_gpu:.1e}, sparse vs plain DIA {e_dia:.1e}",
          flush=True)
    assert e_gpu < 5e-6 and e_dia > 10 * e_gpu
