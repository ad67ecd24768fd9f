"""Pinning of the oracle's 2DIR response (CPU).

src/calc_2DIR.c needs MPI and is not built here, so the oracle's 2DIR
composition is checked against closed forms of uncoupled anharmonic sites
(derived in twodir_closed.py): each pathway alone, the cancellation of the
cross-site terms in their sum (the two-level result per site), fixed and
per-frame anharmonicities with overtone dipoles, lifetime damping, and
T1 != T3.  The propagators and dipole operators it is built from are pinned
to the reference's golden vectors in test_oracle_pin.py.
"""
import numpy as np
import pytest

from nise_2017_amd import abi, synth

import twodir_closed as CF


def _non(N, T1, T3, sr, begin, end, H, technique="2DIR", anh=16.0, lifetime=0.0, ts=1):
    lo, hi = synth.shifted_center(H, N)
    return abi.make_non(singles=N, tmax1=T1, tmax2=0, tmax3=T3, sample=sr, begin=begin, end=end,
                        ts=ts, propagation=1, deltat=2.0, min1=lo, max1=hi, couplingcut=0.0,
                        technique=technique, anharmonicity=anh, lifetime=lifetime)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("technique,paths", [("GBIR", ("GB",)), ("SEIR", ("SE",)),
                                              ("EAIR", ("EA",)), ("noEAIR", ("GB", "SE")),
                                              ("2DIR", ("GB", "SE", "EA"))])
def test_pathways_uncoupled(orc, technique, paths):
    N, T1, T3 = 3, 5, 6
    H, mu, anh, over = synth.twodir_system(N=N, nframes=60, coupled=False, seed=11)
    non = _non(N, T1, T3, 4, 1, 4, H, technique=technique, lifetime=300.0)
    R = orc.twodir(non, H, mu)
    ref = CF.response(non, H, mu, pathways=paths)
    assert _rel(R, ref) < 2e-5


def test_two_level_cancellation(orc):
    """All pathways together: cross-site terms cancel (per-site result)."""
    N, T1, T3 = 4, 6, 4
    H, mu, anh, over = synth.twodir_system(N=N, nframes=60, coupled=False, seed=12)
    non = _non(N, T1, T3, 5, 0, 3, H)
    R = orc.twodir(non, H, mu)
    assert _rel(R, CF.two_level(non, H, mu)) < 2e-5
    # and the closed form's own pathway sum agrees with the per-site form
    assert _rel(CF.response(non, H, mu), CF.two_level(non, H, mu)) < 1e-12


def test_anharmonicity_records(orc):
    """anharmonicity 0: per-frame Delta and overtone dipoles from records."""
    N, T1, T3 = 3, 4, 7
    H, mu, anh, over = synth.twodir_system(N=N, nframes=60, coupled=False, seed=13)
    non = _non(N, T1, T3, 3, 2, 5, H, anh=0.0, lifetime=500.0)
    R = orc.twodir(non, H, mu, A_rec=anh, over_rec=over)
    assert _rel(R, CF.two_level(non, H, mu, anh, over)) < 2e-5
    assert _rel(R, CF.response(non, H, mu, anh, over)) < 2e-5


def test_single_site_textbook(orc):
    """N = 1: S_I = -2 O e^{i(Phi1 - Phi3)} (1 - e^{i Delta f t3})."""
    H, mu, anh, over = synth.twodir_system(N=1, nframes=40, coupled=False, seed=14)
    non = _non(1, 6, 6, 2, 0, 4, H, anh=20.0)
    R = orc.twodir(non, H, mu)
    assert _rel(R, CF.two_level(non, H, mu)) < 2e-5
    # cross polarisation vanishes only for parallel dipoles: here they rotate
    assert np.abs(R[8:]).max() > 1e-3 * np.abs(R[:4]).max()


def test_closed_form_detects_frame_shift(orc):
    """A closed form evaluated one frame late must fail the same bound."""
    N, T1, T3 = 2, 5, 5
    H, mu, anh, over = synth.twodir_system(N=N, nframes=60, coupled=False, seed=15)
    non = _non(N, T1, T3, 4, 1, 3, H)
    R = orc.twodir(non, H, mu)
    Hs = np.roll(H, -1, axis=0)
    mus = np.roll(mu, -1, axis=0)
    assert _rel(R, CF.two_level(non, Hs, mus)) > 1e-3


@pytest.mark.parametrize("technique,paths", [("GBUVvis", ("GB",)), ("SEUVvis", ("SE",)),
                                              ("EAUVvis", ("EA",)), ("noEAUVvis", ("GB", "SE")),
                                              ("2DUVvis", ("GB", "SE", "EA"))])
def test_two_level_sites_uncoupled(orc, technique, paths):
    """calc_2DES (2DUVvis): uncoupled two-level sites, each pathway alone and
    together; the overtone terms are absent, whatever the anharmonicity."""
    N, T1, T3 = 3, 5, 6
    H, mu, anh, over = synth.twodir_system(N=N, nframes=60, coupled=False, seed=21)
    non = _non(N, T1, T3, 4, 1, 4, H, technique=technique, lifetime=300.0, anh=16.0)
    R = orc.twodir(non, H, mu)
    assert _rel(R, CF.response(non, H, mu, pathways=paths, es=True)) < 2e-5
    if technique == "2DUVvis":
        assert _rel(R, CF.two_level(non, H, mu, es=True)) < 2e-5
        # the anharmonic-oscillator result differs: the overtone pathway matters
        assert _rel(R, CF.two_level(non, H, mu)) > 1e-2
