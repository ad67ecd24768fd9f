"""Synthetic NISE trajectories in the reference's binary layout (GROBIN).

Every frame is one int32 header followed by float32 payload
(doc/input.tex:58-90, src/read_trajectory.c:138-176,357-377):
  Full Hamiltonian  : N(N+1)/2 floats, packed upper triangle, row major
  Coupling-type H   : N diagonal floats (static couplings in one Full record)
  Dipole            : 3N floats, x block, y block, z block
Arrays returned here are float32 records of shape (frames, stride) whose
first column holds the header word (the frame number, stored as int32 bits).

Site energies follow the AR(1) process x_{t+1} = a x_t + sqrt(1-a^2) s xi,
a = exp(-dt/tau), started from its stationary distribution, the same form as
the reference's tutorial generator (example/tutorial/src/stochastic.c:29-41).
"""
import numpy as np

SEED0 = 20261202


def sindex(a, b, N):
    if a > b:
        a, b = b, a
    return b + a * (2 * N - a - 1) // 2


def tri_index_arrays(N):
    """(rows, cols) of the packed upper triangle in storage order."""
    a, b = np.triu_indices(N)
    return a, b


def ar1(rng, nframes, nsites, mean, sd, tau, dt):
    a = np.exp(-dt / tau)
    b = np.sqrt(1.0 - a * a)
    x = np.empty((nframes, nsites), np.float64)
    x[0] = rng.normal(0.0, sd, nsites)
    noise = rng.normal(0.0, sd, (nframes - 1, nsites))
    for t in range(1, nframes):
        x[t] = a * x[t - 1] + b * noise[t - 1]
    return (x + mean).astype(np.float32)


def unit_vectors(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def headers(nframes):
    return np.arange(nframes, dtype=np.int32).view(np.float32)


def full_records(diag, coup_tri):
    """Full Hamiltonian records from per-frame diagonals (T, N) and per-frame
    or static strict-upper couplings (T, N(N-1)/2) or (N(N-1)/2,)."""
    T, N = diag.shape
    nn2 = N * (N + 1) // 2
    rec = np.empty((T, 1 + nn2), np.float32)
    rec[:, 0] = headers(T)
    a, b = tri_index_arrays(N)
    is_diag = a == b
    pos = 1 + np.arange(nn2)
    rec[:, pos[is_diag]] = diag
    rec[:, pos[~is_diag]] = coup_tri
    return rec


def dipole_records(mu_site, nframes):
    """Static per-site dipoles (N, 3) repeated over frames: (T, 1 + 3N)."""
    N = mu_site.shape[0]
    rec = np.empty((nframes, 1 + 3 * N), np.float32)
    rec[:, 0] = headers(nframes)
    rec[:, 1:] = np.concatenate([mu_site[:, 0], mu_site[:, 1], mu_site[:, 2]])[None, :]
    return rec


def chain_couplings(N, J):
    """Strict upper triangle (storage order) with J[i] on (i, i+1)."""
    a, b = np.triu_indices(N, 1)
    out = np.zeros(a.size, np.float32)
    nn = b == a + 1
    out[nn] = J[a[nn]]
    return out


def dipole_coupling(pos, mu, prefactor, cut):
    """J_ij = prefactor * (mu_i.mu_j - 3 (mu_i.r)(mu_j.r)) / r^3 (unit r in
    lattice units), |J| < cut -> 0.  Strict upper triangle, storage order."""
    N = pos.shape[0]
    a, b = np.triu_indices(N, 1)
    r = pos[b] - pos[a]
    d = np.linalg.norm(r, axis=1)
    rh = r / d[:, None]
    kappa = np.einsum("ij,ij->i", mu[a], mu[b]) - 3 * np.einsum("ij,ij->i", mu[a], rh) * np.einsum(
        "ij,ij->i", mu[b], rh)
    J = prefactor * kappa / d ** 3
    J[np.abs(J) < cut] = 0.0
    return J.astype(np.float32)


# ---------------------------------------------------------------------------
# The benchmark configurations
# ---------------------------------------------------------------------------

def chain_absorption(N=16, nframes=4200, seed=SEED0, dt=2.0, mean=1650.0, sd=10.0, tau=100.0,
                     jmean=-5.0, jsd=2.0, jfluct=1.0):
    """N-site chain, Full type: AR(1) site energies, nearest-neighbour
    couplings static N(jmean, jsd) plus AR(1) fluctuation of sd jfluct,
    fixed random unit dipoles."""
    rng = np.random.default_rng(seed)
    diag = ar1(rng, nframes, N, mean, sd, tau, dt)
    j0 = rng.normal(jmean, jsd, N - 1)
    jf = ar1(rng, nframes, N - 1, 0.0, jfluct, tau, dt) if jfluct > 0 else np.zeros((nframes, N - 1), np.float32)
    a, b = np.triu_indices(N, 1)
    nn = np.nonzero(b == a + 1)[0]
    coup = np.zeros((nframes, a.size), np.float32)
    coup[:, nn] = (j0[None, :] + jf).astype(np.float32)
    H = full_records(diag, coup)
    mu = dipole_records(unit_vectors(rng, N), nframes)
    return H, mu


def lattice_static(L=16, seed=SEED0 + 1, prefactor=8.0, cut=0.1):
    """L x L square lattice (spacing 1) with random unit dipoles: site
    positions, dipoles, and the static dipole-dipole couplings."""
    rng = np.random.default_rng(seed)
    N = L * L
    pos = np.array([[i % L, i // L, 0.0] for i in range(N)], np.float64)
    mu = unit_vectors(rng, N)
    coup = dipole_coupling(pos, mu, prefactor, cut)
    return rng, mu, coup


def lattice_diag(rng, nframes, N, dt=2.0, mean=1650.0, sd=10.0, tau=100.0):
    return ar1(rng, nframes, N, mean, sd, tau, dt)


def lattice_absorption(L=16, nframes=50000, seed=SEED0 + 1, dt=2.0):
    """Full-type records for the L x L dipole lattice (host memory)."""
    rng, mu, coup = lattice_static(L, seed)
    N = L * L
    diag = lattice_diag(rng, nframes, N, dt)
    return full_records(diag, coup), dipole_records(mu, nframes)


def shifted_center(H_full, N):
    """(min1, max1) window centred on the mean site energy."""
    a, b = np.triu_indices(N)
    d = H_full[:, 1:][:, a == b]
    c = float(np.mean(d))
    return c - 100.0, c + 100.0


def twodir_system(N=6, nframes=200, seed=SEED0 + 2, dt=2.0, mean=1650.0, sd=10.0, tau=100.0,
                  coupled=True, jmean=-5.0, jsd=2.0, dip_fluct=0.2, anh_mean=16.0, anh_sd=2.0):
    """Records for a 2DIR run: Full Hamiltonian (chain couplings, or none
    when coupled is False), fluctuating dipoles (unit vectors plus AR(1)
    noise of sd dip_fluct per component), per-frame anharmonicities
    (1 + N words) and overtone dipoles (1 + 3N words, sqrt(2) mu plus
    noise) for runs with anharmonicity 0."""
    rng = np.random.default_rng(seed)
    diag = ar1(rng, nframes, N, mean, sd, tau, dt)
    a, b = np.triu_indices(N, 1)
    coup = np.zeros((nframes, a.size), np.float32)
    if coupled and N > 1:
        nn = np.nonzero(b == a + 1)[0]
        j0 = rng.normal(jmean, jsd, N - 1)
        jf = ar1(rng, nframes, N - 1, 0.0, 1.0, tau, dt)
        coup[:, nn] = (j0[None, :] + jf).astype(np.float32)
    H = full_records(diag, coup)
    u = unit_vectors(rng, N)
    noise = ar1(rng, nframes, 3 * N, 0.0, dip_fluct, tau, dt).reshape(nframes, 3, N)
    mu_xyz = (u.T[None, :, :] + noise).astype(np.float32)  # (T, 3, N)
    mu = np.empty((nframes, 1 + 3 * N), np.float32)
    mu[:, 0] = headers(nframes)
    mu[:, 1:] = mu_xyz.reshape(nframes, 3 * N)
    anh = np.empty((nframes, 1 + N), np.float32)
    anh[:, 0] = headers(nframes)
    anh[:, 1:] = ar1(rng, nframes, N, anh_mean, anh_sd, tau, dt)
    over = np.empty((nframes, 1 + 3 * N), np.float32)
    over[:, 0] = headers(nframes)
    on = ar1(rng, nframes, 3 * N, 0.0, dip_fluct, tau, dt)
    over[:, 1:] = (np.sqrt(2.0) * mu_xyz.reshape(nframes, 3 * N) + on).astype(np.float32)
    return H, mu, anh, over


def geometric_system(N, nframes, seed, dt=1.0, mean=15000.0, sd=100.0, tau=50.0, spacing=5.0,
                     prefactor=1000.0, cut=0.5):
    """Coupling-type records for N sites uniform in a cube at density one per
    spacing^3 (Angstrom), unit transition dipoles with random orientation,
    point-dipole couplings J = prefactor * kappa / r^3 (cm^-1, r in
    Angstrom; the default gives 8 cm^-1 at 5 Angstrom for kappa = 1),
    |J| < cut dropped; AR(1) site energies (mean, sd, tau fs).
    Returns (H diagonal records (T, 1 + N), coupling record (1 + N(N+1)/2,),
    dipole record (1, 1 + 3N))."""
    rng = np.random.default_rng(seed)
    side = spacing * N ** (1.0 / 3.0)
    pos = rng.uniform(0.0, side, (N, 3))
    mu = unit_vectors(rng, N)
    J = np.zeros(N * (N - 1) // 2, np.float32)
    if N > 1:
        a, b = np.triu_indices(N, 1)
        r = pos[b] - pos[a]
        d = np.maximum(np.linalg.norm(r, axis=1), 1.0)
        rh = r / d[:, None]
        kappa = np.einsum("ij,ij->i", mu[a], mu[b]) - 3 * np.einsum("ij,ij->i", mu[a], rh) * \
            np.einsum("ij,ij->i", mu[b], rh)
        J = (prefactor * kappa / d ** 3).astype(np.float32)
        J[np.abs(J) < cut] = 0.0
    diag = ar1(rng, nframes, N, mean, sd, tau, dt)
    H = np.empty((nframes, 1 + N), np.float32)
    H[:, 0] = headers(nframes)
    H[:, 1:] = diag
    nn2 = N * (N + 1) // 2
    Crec = np.zeros(1 + nn2, np.float32)
    ai, bi = np.triu_indices(N)
    Crec[1 + np.nonzero(ai != bi)[0]] = J
    return H, Crec, dipole_records(mu, 1)


def ring_2dir(N=32, nframes=4200, seed=SEED0 + 2, dt=2.0, mean=1650.0, sd=12.0, tau=100.0,
              jmean=-6.0, jsd=2.0):
    """N-site ring, Full type: AR(1) site energies, static nearest-neighbour
    couplings N(jmean, jsd) including the (0, N-1) bond, fixed random unit
    dipoles.  Returns (H, mu) records."""
    rng = np.random.default_rng(seed)
    diag = ar1(rng, nframes, N, mean, sd, tau, dt)
    a, b = np.triu_indices(N, 1)
    coup = np.zeros(a.size, np.float32)
    J = rng.normal(jmean, jsd, N).astype(np.float32)
    for i in range(N):
        x, y = sorted((i, (i + 1) % N))
        if x != y:
            coup[np.nonzero((a == x) & (b == y))[0]] = J[i]
    H = full_records(diag, coup)
    mu = dipole_records(unit_vectors(rng, N), nframes)
    return H, mu


def onfly_system(N, nframes, seed, kind=1, box=None, spacing=5.0, mu_debye=0.5, arm=0.6, dt=2.0,
                 mean=1650.0, sd=10.0, tau=100.0, jitter=0.05, wobble=0.02):
    """Records for HamiltonianType TransitionDipole (kind 1) or
    ExtendedDipole (kind 2): site energies (T, 1 + N) as AR(1), transition
    dipoles (T, 1 + 3N) of magnitude mu_debye wobbling about random
    directions, positions (Angstrom) on a jittered cubic grid of the given
    spacing moving by small random steps.  ExtendedDipole has two position
    records per frame (2T, 1 + 3N): the charges at centre -/+ arm/2 along
    the dipole.  box: cubic edge for the minimum image (positions wrapped
    into [0, box)), giving box records (T, 4), or None.
    Returns (H, mu, pos, box_records)."""
    rng = np.random.default_rng(seed)
    n = int(np.ceil(N ** (1.0 / 3.0)))
    grid = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3)
    centre = (grid[:N] * spacing + rng.normal(0, 0.3, (N, 3))).astype(np.float64)
    axis = unit_vectors(rng, N)
    H = np.empty((nframes, 1 + N), np.float32)
    H[:, 0] = headers(nframes)
    H[:, 1:] = ar1(rng, nframes, N, mean, sd, tau, dt)
    mu = np.empty((nframes, 1 + 3 * N), np.float32)
    mu[:, 0] = headers(nframes)
    per = 1 if kind == 1 else 2
    pos = np.empty((per * nframes, 1 + 3 * N), np.float32)
    pos[:, 0] = headers(per * nframes)
    for f in range(nframes):
        centre += rng.normal(0, jitter, (N, 3))
        axis += rng.normal(0, wobble, (N, 3))
        axis /= np.linalg.norm(axis, axis=1)[:, None]
        m = mu_debye * axis
        mu[f, 1:] = np.concatenate([m[:, 0], m[:, 1], m[:, 2]])
        c = np.mod(centre, box) if box else centre
        if kind == 1:
            pos[f, 1:] = np.concatenate([c[:, 0], c[:, 1], c[:, 2]])
        else:
            for k, sgn in enumerate((-0.5, 0.5)):
                p = c + sgn * arm * axis
                if box:
                    p = np.mod(p, box)
                pos[2 * f + k, 1:] = np.concatenate([p[:, 0], p[:, 1], p[:, 2]])
    boxes = None
    if box:
        boxes = np.empty((nframes, 4), np.float32)
        boxes[:, 0] = headers(nframes)
        boxes[:, 1:] = box
    return H, mu, pos, boxes


def uneven_batch(nsys=96, seed=SEED0 + 4, samples=100, tmax=512, samplerate=10, sys_seed=SEED0 + 4000,
                 full_samples=100):
    """The uneven-size batch: nsys systems with N = round(exp U[ln 8, ln
    1024]), each a geometric_system (sites uniform in a cube at one per
    (5 A)^3, couplings 8 cm^-1 at 5 A for kappa = 1 dropped below 0.5
    cm^-1, AR(1) energies 15000 / 100 cm^-1 / 50 fs, dt 1 fs); system k
    draws from seed sys_seed + k.  The trajectories are generated for
    full_samples start points every samplerate frames over tmax steps, and
    the window (min1, max1) is centred on their mean; the records returned
    are cut to what `samples` start points need, so a short run is the
    prefix of the full one.  N <= 64 runs on the exact exponential
    (propagation 0), larger N on the Trotter scheme.
    Returns a list of dicts: N, propagation, H (T, 1 + N), C (coupling
    record), mu (dipole record), min1, max1."""
    rng = np.random.default_rng(seed)
    Ns = np.round(np.exp(rng.uniform(np.log(8), np.log(1024), nsys))).astype(int)
    nfr = (full_samples - 1) * samplerate + tmax + 3
    need = (samples - 1) * samplerate + tmax + 3
    out = []
    for k, N in enumerate(Ns):
        H, Cr, mu = geometric_system(int(N), nfr, seed=sys_seed + k)
        mean = float(H[:, 1:].mean())
        out.append(dict(N=int(N), propagation=0 if N <= 64 else 1, H=np.ascontiguousarray(H[:need]), C=Cr,
                        mu=mu, min1=mean - 600.0, max1=mean + 600.0))
    return out
