"""Closed-form 2DIR response of uncoupled anharmonic sites.

Without couplings every site evolves by its own phase: a singles amplitude
on site a picks up exp(-i f e_a(t)) per frame (e_a = H_aa - shift,
f = 2 pi c dt), and a two-exciton amplitude on |ab> picks up
exp(-i f (e_a + e_b - delta_ab Delta_a(t))).  The three pathways of the
third-order response then reduce to sums over site pairs (a, b):

  GB_I  = -sum_ab O(A_a, B_a, C_b, D_b) conj(z1_a) z3_b
  GB_II = -sum_ab O(A_a, B_a, C_b, D_b) z1_a z3_b
  SE_I  = -sum_ab O(A_a, B_b, C_a, D_b) conj(z1_a) z3_b
  SE_II = -sum_ab O(A_a, B_b, C_b, D_a) z1_a z3_a
  EA_I  = sum_{c!=a} [O(A_c, B_c, C_a, D_a) + O(A_c, B_a, C_c, D_a)] conj(z1_c) z3_a
          + sum_c O(A_c, B_c, C'_c, D'_c) conj(z1_c) z3_c zD_c
  EA_II = sum_{c!=a} [O(A_c, B_c, C_a, D_a) z1_c z3_a + O(A_a, B_c, C_c, D_a) z1_a z3_a]
          + sum_c O(A_c, B_c, C'_c, D'_c) z1_c z3_c zD_c

with A = mu(tj - t1), B = mu(tj), C = mu(tk), D = mu(tl) per site, C', D'
the overtone dipoles (sqrt 2 mu for a fixed anharmonicity), z1 =
exp(-i Phi_1), z3 = exp(-i Phi_3), zD = exp(+i f sum Delta), and O the
isotropic orientational average of the four dipoles for the parallel,
perpendicular and cross lab polarisations:

  par = [(AB)(CD) + (AC)(BD) + (AD)(BC)] / 15
  per = [4 (AB)(CD) - (AC)(BD) - (AD)(BC)] / 30
  cro = 2.5 [(AC)(BD) - (AD)(BC)] / 30

In the sum of all three pathways the cross-site terms cancel and each site
contributes the two-level result -2 O e^{...} (1 - e^{i Delta t3}) of an
anharmonic oscillator.  Signals are stored as rrX = Im S_X, riX = Re S_X.

es=True gives the same sums for two-level sites (calc_2DES, 2DUVvis): no
site holds two excitations, so the overtone terms are absent; the states
|ab>, a != b, have the energy e_a + e_b and the same mu_ef as above, so the
cross-site terms still cancel and each site contributes -2 O e^{...}.
"""
import numpy as np

ICM2IFS = 2.99792458e-5


def orient(A, B, Cv, D):
    ab, cd, ac, bd, ad, bc = A @ B, Cv @ D, A @ Cv, B @ D, A @ D, B @ Cv
    return np.array([(ab * cd + ac * bd + ad * bc) / 15.0,
                     (4 * ab * cd - ac * bd - ad * bc) / 30.0,
                     2.5 * (ac * bd - ad * bc) / 30.0])


def response(non, H, mu, anh=None, over=None, pathways=("GB", "SE", "EA"), es=False):
    """(12, tmax3, tmax1) float64 planes for uncoupled sites (Full records)."""
    N, T1, T3 = non.singles, non.tmax1, non.tmax3
    shift = np.float32((np.float32(non.max1) + np.float32(non.min1)) / np.float32(2))
    f = float(np.float32(non.deltat)) * 2 * np.pi * ICM2IFS
    a_idx, b_idx = np.triu_indices(N)
    diag = H[:, 1:][:, a_idx == b_idx]
    e = (diag - shift).astype(np.float32).astype(np.float64)  # (T, N)
    mu3 = mu[:, 1:].reshape(-1, 3, N).transpose(0, 2, 1).astype(np.float64)  # (T, N, 3)
    if es:
        ov3, dl = np.zeros_like(mu3), np.zeros((H.shape[0], N))
    elif non.anharmonicity != 0:
        ov3 = np.float64(np.float32(np.sqrt(2.0))) * mu3
        dl = np.full((H.shape[0], N), float(np.float32(non.anharmonicity)))
    else:
        ov3 = over[:, 1:].reshape(-1, 3, N).transpose(0, 2, 1).astype(np.float64)
        dl = anh[:, 1:].astype(np.float64)
    lt = np.ones((T3, T1))
    if non.lifetime > 0:
        t3, t1 = np.meshgrid(np.arange(T3), np.arange(T1), indexing="ij")
        lt = np.exp(-(t1 + t3) * float(np.float32(non.deltat)) / (2 * float(np.float32(non.lifetime))))
    SI = np.zeros((3, T3, T1), complex)
    SII = np.zeros((3, T3, T1), complex)
    end = non.end if non.end else (H.shape[0] - T1 - non.tmax2 - T3 - 1) // non.sample + 1
    for s in range(non.begin, end):
        tj = s * non.sample + T1
        tk = tj
        for t1 in range(T1):
            z1 = np.exp(-1j * f * e[tj - t1:tj].sum(axis=0))
            A = mu3[tj - t1]
            B = mu3[tj]
            Cv = mu3[tk]
            Cp = ov3[tk]
            for t3 in range(T3):
                tl = tk + t3
                z3 = np.exp(-1j * f * e[tk:tl].sum(axis=0))
                zD = np.exp(1j * f * dl[tk:tl].sum(axis=0))
                D = mu3[tl]
                Dp = ov3[tl]
                sI = np.zeros(3, complex)
                sII = np.zeros(3, complex)
                for a in range(N):
                    for b in range(N):
                        if "GB" in pathways:
                            o = orient(A[a], B[a], Cv[b], D[b])
                            sI -= o * np.conj(z1[a]) * z3[b]
                            sII -= o * z1[a] * z3[b]
                        if "SE" in pathways:
                            sI -= orient(A[a], B[b], Cv[a], D[b]) * np.conj(z1[a]) * z3[b]
                            sII -= orient(A[a], B[b], Cv[b], D[a]) * z1[a] * z3[a]
                        if "EA" in pathways and a != b:
                            c = b  # c != a
                            sI += (orient(A[c], B[c], Cv[a], D[a]) +
                                   orient(A[c], B[a], Cv[c], D[a])) * np.conj(z1[c]) * z3[a]
                            sII += (orient(A[c], B[c], Cv[a], D[a]) * z1[c] * z3[a] +
                                    orient(A[a], B[c], Cv[c], D[a]) * z1[a] * z3[a])
                    if "EA" in pathways and not es:
                        o = orient(A[a], B[a], Cp[a], Dp[a])
                        sI += o * np.conj(z1[a]) * z3[a] * zD[a]
                        sII += o * z1[a] * z3[a] * zD[a]
                SI[:, t3, t1] += sI * lt[t3, t1]
                SII[:, t3, t1] += sII * lt[t3, t1]
    R = np.zeros((12, T3, T1))
    for p in range(3):
        R[4 * p + 0], R[4 * p + 1] = SI[p].imag, SI[p].real
        R[4 * p + 2], R[4 * p + 3] = SII[p].imag, SII[p].real
    return R


def two_level(non, H, mu, anh=None, over=None, es=False):
    """The sum of all pathways written per site, without cross terms."""
    N, T1, T3 = non.singles, non.tmax1, non.tmax3
    shift = np.float32((np.float32(non.max1) + np.float32(non.min1)) / np.float32(2))
    f = float(np.float32(non.deltat)) * 2 * np.pi * ICM2IFS
    a_idx, b_idx = np.triu_indices(N)
    e = (H[:, 1:][:, a_idx == b_idx] - shift).astype(np.float32).astype(np.float64)
    mu3 = mu[:, 1:].reshape(-1, 3, N).transpose(0, 2, 1).astype(np.float64)
    if es:
        ov3, dl = np.zeros_like(mu3), np.zeros((H.shape[0], N))
    elif non.anharmonicity != 0:
        ov3 = np.float64(np.float32(np.sqrt(2.0))) * mu3
        dl = np.full((H.shape[0], N), float(np.float32(non.anharmonicity)))
    else:
        ov3 = over[:, 1:].reshape(-1, 3, N).transpose(0, 2, 1).astype(np.float64)
        dl = anh[:, 1:].astype(np.float64)
    lt = np.ones((T3, T1))
    if non.lifetime > 0:
        t3, t1 = np.meshgrid(np.arange(T3), np.arange(T1), indexing="ij")
        lt = np.exp(-(t1 + t3) * float(np.float32(non.deltat)) / (2 * float(np.float32(non.lifetime))))
    R = np.zeros((12, T3, T1))
    end = non.end if non.end else (H.shape[0] - T1 - non.tmax2 - T3 - 1) // non.sample + 1
    for s in range(non.begin, end):
        tj = tk = s * non.sample + T1
        for t1 in range(T1):
            for t3 in range(T3):
                tl = tk + t3
                for a in range(N):
                    p1 = f * e[tj - t1:tj, a].sum()
                    p3 = f * e[tk:tl, a].sum()
                    pd = f * dl[tk:tl, a].sum()
                    o = orient(mu3[tj - t1, a], mu3[tj, a], mu3[tk, a], mu3[tl, a])
                    oe = orient(mu3[tj - t1, a], mu3[tj, a], ov3[tk, a], ov3[tl, a])
                    sI = (-2 * o + oe * np.exp(1j * pd)) * np.exp(1j * (p1 - p3)) * lt[t3, t1]
                    sII = (-2 * o + oe * np.exp(1j * pd)) * np.exp(-1j * (p1 + p3)) * lt[t3, t1]
                    for p in range(3):
                        R[4 * p + 0, t3, t1] += sI[p].imag
                        R[4 * p + 1, t3, t1] += sI[p].real
                        R[4 * p + 2, t3, t1] += sII[p].imag
                        R[4 * p + 3, t3, t1] += sII[p].real
    return R
