"""The 2D cases of tests/golden/reference_2d.npz: input decks, trajectory
records and the R*.dat planes the reference's own calc_2DIR / calc_2DES
wrote for them (tests/golden/make_golden_2d.py)."""
import json
import os

import numpy as np

from nise_2017_amd import abi

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_2d.npz")
FILES = ("RparI.dat", "RparII.dat", "RperI.dat", "RperII.dat", "RcroI.dat", "RcroII.dat")
ICM2IFS = np.float32(2.99792458e-5)
TWOPI = np.float32(2 * 3.14159265359)

_data = None


def data():
    global _data
    if _data is None:
        _data = dict(np.load(GOLD, allow_pickle=False))
    return _data


def names():
    return json.loads(str(data()["cases"]))


def case(name):
    """(non, records dict, deck dict, planes (6, rows, 5)) of one case.

    non is what readInput (src/readinput.c:24-289) leaves for the deck:
    tmax3 = tmax1, propagation 0 unless 'Coupling', the threshold of the
    sparse scheme multiplied by (dt icm2ifs 2 pi / Trotter)^2 in FP32."""
    k = names().index(name)
    d = data()
    deck = json.loads(str(d[f"c{k}_deck"]))
    rec = {key[len(f"c{k}_"):]: d[key] for key in d if key.startswith(f"c{k}_") and key not in
           (f"c{k}_deck", f"c{k}_R")}
    T1, T2, _ = deck["RunTimes"]
    prop = 1 if deck.get("Propagation") == "Coupling" else 0
    ts = int(deck.get("Trotter", 5))
    thres = np.float32(deck.get("Threshold", 0.0))
    if prop == 0:
        g = np.float32(np.float32(np.float32(np.float32(deck["Timestep"]) * ICM2IFS) * TWOPI) / np.float32(ts))
        thres = np.float32(thres * np.float32(g * g))
    kw = dict(singles=int(deck["Singles"]), tmax1=int(T1), tmax2=int(T2), tmax3=int(T1),
              sample=int(deck["Samplerate"]), begin=int(deck.get("BeginPoint", 0)), end=int(deck["EndPoint"]),
              ts=ts, propagation=prop, deltat=float(deck["Timestep"]),
              lifetime=float(deck.get("Lifetime", 0.0)), anharmonicity=float(deck.get("Anharmonicity", 0.0)),
              couplingcut=float(deck.get("Couplingcut", 0.0)), technique=deck["Technique"],
              hamiltonian=deck.get("HamiltonianType", "Full"), min1=float(deck["MinFrequencies"][0]),
              max1=float(deck["MaxFrequencies"][0]), length=int(deck["Length"]), fft=int(deck.get("FFT", 0)))
    non = abi.make_non(**kw)
    non.thres = thres
    non.min2, non.min3 = non.min1, non.min1
    non.max2, non.max3 = non.max1, non.max1
    return non, rec, deck, d[f"c{k}_R"]


def print2d(non, Rsum, samples):
    """src/MPI_subs.c:20-34 applied to 12 summed planes: (6, T1*T3, 5)."""
    out = []
    for p in range(3):
        for k in range(2):
            re = Rsum[4 * p + 2 * k].astype(np.float32) / np.float32(samples)
            im = Rsum[4 * p + 2 * k + 1].astype(np.float32) / np.float32(samples)
            re[0] /= 2
            im[0] /= 2
            rows = []
            for t1 in range(non.tmax1):
                for t3 in range(non.tmax3):
                    rows.append([t1 * non.deltat, non.tmax2 * non.deltat, t3 * non.deltat, re[t3, t1], im[t3, t1]])
            out.append(rows)
    return np.array(out, np.float64)


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
