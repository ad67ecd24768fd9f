"""The autotuner (nise_abs_plan_autotune) on a small lattice: it times every
candidate (row conveyor with 1..64 rows, slot-conveyor shapes), keeps the
fastest, and the plan's result afterwards matches the oracle within 1e-5
(the batched path's bar).  Each candidate set explicitly gives the same
response within 1e-6 (the propagation is the reference's FP32 order in
every shape; only the grouping of the FP64 partial sums differs)."""
import numpy as np
import pytest

from nise_2017_amd import abi, synth

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_autotune_keeps_the_fastest_and_stays_exact(orc):
    L, T, sr, ns = 6, 32, 5, 40
    nfr = (ns - 1) * sr + T + 4
    H, mu = synth.lattice_absorption(L=L, nframes=nfr, seed=901)
    N = L * L
    lo, hi = synth.shifted_center(H, N)
    non = abi.make_non(singles=N, tmax1=T, sample=sr, begin=0, end=ns, ts=1, propagation=1, deltat=2.0,
                       min1=lo, max1=hi)
    non.shifte = np.float32((np.float32(non.max1) + np.float32(non.min1)) / np.float32(2))
    plan = abi.AbsPlan(non, H, H.shape[1], mu, mu.shape[1], nfr)
    try:
        cands = plan.autotune()
        print(f"[autotune] candidates {cands}, chosen {plan.info()}", flush=True)
        assert len(cands) >= 2 and all(ms > 0 for _, ms in cands)
        info = plan.info()
        if info["rows"] > 0:
            best = min(cands, key=lambda c: c[1])
            assert best[0] == info["rows"]
        re, im = plan.execute()
        o_re, o_im = orc.absorption(non, H, mu)
        assert rel(re + 1j * im, o_re + 1j * o_im) < 1e-5
        for code, _ in cands:
            if code > 0:
                plan.set_rows(code)
            else:
                c = -code
                vp, spw, blk = c // 10000, (c // 1000) % 10, c % 1000
                plan.set_shape(vp, spw, blk if blk < 900 else -(blk - 900))
            r1, i1 = plan.execute()
            assert rel(r1 + 1j * i1, re + 1j * im) < 1e-6, code
    finally:
        plan.close()
