#!/usr/bin/env python3
"""Benchmark: batched linear absorption on one MI355X.

Workload (the headline configuration): N=256 sites on a 16x16 square lattice,
dipole-dipole couplings J = 8 cm^-1 (a/r)^3 kappa truncated below 0.1 cm^-1,
AR(1) site energies (mean 1650, sd 10 cm^-1, correlation 100 fs), 50,000
FP32 Full-Hamiltonian frames at dt = 2 fs, tmax = 256 steps, 4000 start
samples spaced 10 frames apart, 3 polarisations, Trotter 1.  Synthetic data,
resident in HBM before the timed region.

One step = one complete absorption run: per-frame rotation/phase tables,
the frame-major conveyor kernel and the response reduction, replayed as a
captured hipGraph.  value = site-steps per second, a site-step being one
amplitude of one (sample, polarisation) vector advanced by one frame.

Also reported on the same line: the roofline of the dominant (conveyor)
kernel from HIP events on its own stream, the plain one-sample-per-launch
GPU kernel on a subset of samples, and the CPU restatement of the reference
FP32 path timed on this host's cores (baseline only).
"""
import argparse
import contextlib
import gc
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


@contextlib.contextmanager
def no_gc():
    """Python's collector paused over a timed region (a collection pause
    between two library calls is the interpreter's, not the GPU path's)."""
    gc.collect()
    gc.disable()
    try:
        yield
    finally:
        gc.enable()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes_per_site_step(N, tmax, sample, full=True):
    """B = 16 + (F_H + F_mu) / (N * V_f), V_f = 3 tmax / Samplerate."""
    F_H = 4 + 4 * (N * (N + 1) // 2 if full else N)
    F_mu = 4 + 12 * N
    V_f = 3.0 * tmax / sample
    return 16.0 + (F_H + F_mu) / (N * V_f)


def build_lattice_on_device(L, nframes, seed, torch):
    import numpy as np

    from nise_2017_amd import synth
    rng, mus, coup = synth.lattice_static(L, seed=seed)
    N = L * L
    nn2 = N * (N + 1) // 2
    diag = synth.lattice_diag(rng, nframes, N)
    a, b = np.triu_indices(N)
    row = np.zeros(1 + nn2, np.float32)
    row[1 + np.nonzero(a != b)[0]] = coup
    H = torch.empty((nframes, 1 + nn2), dtype=torch.float32, device="cuda")
    H.copy_(torch.from_numpy(row).cuda().expand(nframes, -1))
    didx = torch.from_numpy(1 + np.nonzero(a == b)[0]).cuda()
    H[:, didx] = torch.from_numpy(diag).cuda()
    mu = torch.from_numpy(synth.dipole_records(mus, 1)[0]).cuda().expand(nframes, -1).contiguous()
    return H, mu, diag, coup, mus, row


def host_records(diag, row, mus, f_hi):
    import numpy as np

    from nise_2017_amd import synth
    N = diag.shape[1]
    a, b = np.triu_indices(N)
    H = np.repeat(row[None, :], f_hi, axis=0)
    H[:, 1 + np.nonzero(a == b)[0]] = diag[:f_hi]
    return H, synth.dipole_records(mus, f_hi)


def adjoint_leg(args, torch):
    """Gradient of the absorption spectrum for the N=128 chain: 20,000
    frames, tmax 256, 1000 samples every 19 frames, Trotter 1; parameters
    128 site offsets (0) and the coupling scale (1); loss |spec - target|^2
    on the 2048-point grid, target from offsets N(0, 5) and scale 1.1.
    Reports forward and adjoint times per checkpoint interval K, memory
    against storing every state, and the gradient against central finite
    differences of the FP64 oracle (h = 0.01 cm^-1, 1e-4 for the scale)."""
    import numpy as np

    from nise_2017_amd import abi, adjoint, synth
    N, T, sr, ns, nfr = 128, 256, 19, 1000, 20000
    log("adjoint leg: N=128 chain, 1000 samples")
    H, mu = synth.chain_absorption(N=N, nframes=nfr, seed=20261202 + 3)
    need = (ns - 1) * sr + T
    lo, hi = synth.shifted_center(H[:need], N)
    non = abi.make_non(singles=N, tmax1=T, sample=sr, begin=0, end=ns, ts=1, propagation=1,
                       deltat=2.0, min1=lo, max1=hi, fft=2048, lifetime=1300.0)
    non.shifte = np.float32((np.float32(hi) + np.float32(lo)) / np.float32(2))
    Hd, mud = torch.from_numpy(H).cuda(), torch.from_numpy(mu).cuda()
    freq, A_re, A_im = adjoint.spectrum_matrix(non, ns)
    rng = np.random.default_rng(20261202 + 3)
    d_t = rng.normal(0, 5, N).astype(np.float32)
    out = {"workload": "N=128 chain, 20000 frames, tmax 256, 1000 samples every 19 frames, 3 pols",
           "params": N + 1, "grid_points": int(len(freq)), "runs": []}
    gref = None
    for K in [int(k) for k in args.adjoint_every.split(",")] + [1]:
        model = adjoint.AdjointAbsorption(non, Hd, mud, every=K)
        if gref is None:
            St = model.forward(d_t, 1.1)
            target = A_re @ St[0] + A_im @ St[1]
        model.forward(None, 1.0, timed=True)  # warm
        S = model.forward(None, 1.0, timed=True)
        t = model.timing()
        fwd = t["tables_ms"] + t["conveyor_ms"] + t["reduce_ms"]
        r = A_re @ S[0] + A_im @ S[1] - target
        g_re, g_im = 2 * A_re.T @ r, 2 * A_im.T @ r
        model.backward(g_re, g_im)
        gd, gc = model.backward(g_re, g_im)
        adj = model.timing()["adjoint_ms"]
        nck = (T - 1) // K + 1 + (1 if (T - 1) % K else 0)
        out["runs"].append({"K": K, "forward_ms": round(fwd, 3), "adjoint_ms": round(adj, 3),
                            "adjoint_over_forward": round(adj / fwd, 3),
                            "checkpoint_bytes": 3 * ns * nck * N * 8,
                            "all_states_bytes": 3 * ns * T * N * 8})
        log(f"adjoint K={K}: forward {fwd:.2f} ms, adjoint {adj:.2f} ms")
        if gref is None:
            gref = (gd, gc, float(r @ r))
        model.close()
    del Hd, mud
    if args.fd_params > 0 and not args.no_cpu:
        from oracle import oracle
        threads = max(1, min(16, os.cpu_count() or 1))
        gd, gc, L0 = gref

        def L(d, c):
            re, im = oracle.absorption_param_f64(non, H, mu, d, c, nthreads=threads)
            rr = A_re @ re + A_im @ im - target
            return float(rr @ rr)
        idx = list(np.argsort(-np.abs(gd))[:args.fd_params])
        fd = []
        d0 = np.zeros(N)
        for a in idx:
            e = np.zeros(N)
            e[a] = 0.01
            fd.append((L(d0 + e, 1.0) - L(d0 - e, 1.0)) / 0.02)
        fdc = (L(d0, 1.0 + 1e-4) - L(d0, 1.0 - 1e-4)) / 2e-4
        g_sel = np.array([gd[a] for a in idx] + [gc])
        g_fd = np.array(fd + [fdc])
        out["fd_check"] = {"params": [f"delta[{a}]" for a in idx] + ["cscale"],
                           "gpu": [float(x) for x in g_sel], "fd_fp64_oracle": [float(x) for x in g_fd],
                           "rel_err": float(np.linalg.norm(g_sel - g_fd) / np.linalg.norm(g_fd)),
                           "h": {"delta": 0.01, "cscale": 1e-4}}
        log(f"adjoint FD check rel err {out['fd_check']['rel_err']:.2e}")
    return out


def multi_leg(args, torch):
    """Uneven-size batch: 96 systems, N = round(exp U[ln 8, ln 1024]), sites
    uniform in a cube at one per (5 A)^3, point-dipole couplings 8 cm^-1 at
    5 A (kappa = 1) dropped below 0.5 cm^-1, AR(1) energies (15000, 100
    cm^-1, 50 fs), Coupling type, dt 1 fs, 1503 frames; tmax 512, 100
    samples every 10 frames, 3 polarisations; N <= 64 on the exact
    exponential (DIA) path, the rest on the Trotter path.  One call runs
    the per-frame tables, the two launches and the reductions.

    Checked against the oracle: every system of the same batch run on its
    first 2 samples (a second call), and the timed 100-sample results of the
    smallest Trotter systems."""
    import numpy as np

    from nise_2017_amd import abi, synth
    T, sr, ns = 512, 10, 100

    def systems_for(samples, dev):
        specs = synth.uneven_batch(nsys=args.multi_systems, samples=samples, tmax=T, samplerate=sr)
        out = []
        for s in specs:
            non = abi.make_non(singles=s["N"], tmax1=T, sample=sr, begin=0, end=samples, ts=1,
                               propagation=s["propagation"], deltat=1.0, min1=s["min1"], max1=s["max1"],
                               hamiltonian="Coupling", couplingcut=0.5)
            non.shifte = np.float32((np.float32(non.max1) + np.float32(non.min1)) / np.float32(2))
            if dev:
                out.append((non, torch.from_numpy(s["H"]).cuda(), torch.from_numpy(s["mu"]).cuda(),
                            torch.from_numpy(s["C"]).cuda()))
            else:
                out.append((non, s["H"], s["mu"], s["C"]))
        return out

    log(f"multi leg: {args.multi_systems} systems, generating")
    systems = systems_for(ns, True)
    Ns = np.array([s[0].singles for s in systems])
    torch.cuda.synchronize()
    abi.absorption_multi(systems)  # warm
    torch.cuda.synchronize()
    tuned, cands = abi.absorption_multi_tune(systems)
    log(f"multi leg: row counts tuned {[(r, round(ms, 2)) for r, ms in cands]}, kept {tuned}")
    walls = []
    with no_gc():
        for _ in range(5):
            t0 = time.perf_counter()
            out, done = abi.absorption_multi(systems)
            torch.cuda.synchronize()
            walls.append(time.perf_counter() - t0)
    wall = float(np.median(walls))
    site_steps = float(sum(3 * ns * T * int(N) for N in Ns))
    res = {"workload": f"{len(Ns)} systems, N log-uniform in [8, 1024] (sum {int(Ns.sum())}), "
                       f"Coupling type, tmax {T}, {ns} samples, 3 pols, dt 1 fs",
           "systems_dia": int((Ns <= 64).sum()), "systems_trotter": int((Ns > 64).sum()),
           "site_steps": site_steps, "wall_ms": round(wall * 1e3, 3),
           "wall_ms_calls": [round(w * 1e3, 3) for w in walls],
           "rows_tuned": {"kept": tuned, "ms_per_call": {str(r): round(ms, 3) for r, ms in cands}},
           "site_steps_per_s": site_steps / wall,
           "completion_ms_p50": float(np.percentile(done, 50)),
           "completion_ms_p99": float(np.percentile(done, 99)),
           "note": "wall_ms: median host wall time of 5 calls, each with its per-frame tables (the "
                   "matrix exponential for DIA systems); completion times are device-clock times of "
                   "each system's last workgroup in the last call"}
    log(f"multi leg: {res['site_steps_per_s']:.3e} site-steps/s, p50 {res['completion_ms_p50']:.2f} ms, "
        f"p99 {res['completion_ms_p99']:.2f} ms")
    if not args.no_cpu:
        from concurrent.futures import ThreadPoolExecutor

        from oracle import oracle
        threads = max(1, min(16, os.cpu_count() or 1))

        def rel(a, b):
            return float(np.linalg.norm(a - b) / np.linalg.norm(b))

        def check(host_systems, results, idx):
            errs = {}

            def run(k):
                non, H, mu, Cr = host_systems[k]
                o_re, o_im = oracle.absorption(non, H, mu, C_rec=Cr)
                re, im = results[k]
                errs[k] = rel(re + 1j * im, o_re + 1j * o_im)
            cost = lambda k: host_systems[k][0].singles ** (3 if host_systems[k][0].propagation == 0 else 2)  # noqa
            with ThreadPoolExecutor(threads) as ex:
                list(ex.map(run, sorted(idx, key=lambda k: -cost(k))))
            return errs
        t_a = time.perf_counter()
        short_dev = systems_for(2, True)
        short_out, _ = abi.absorption_multi(short_dev)
        del short_dev
        e_all = check(systems_for(2, False), short_out, range(len(Ns)))
        small = [k for k in np.argsort(Ns) if Ns[k] > 64][:4]
        e_full = check(systems_for(ns, False), out, small)
        res["parity"] = {
            "all_systems_2_samples_max_rel_l2": max(e_all.values()),
            "timed_run_smallest_trotter": {f"N={int(Ns[k])}": e_full[k] for k in small},
            "oracle": "FP32 oracle per system (DIA: FP64 eigensystem, FP32 vectors)",
            "seconds": round(time.perf_counter() - t_a, 1)}
        log(f"multi leg parity: all systems (2 samples) max rel L2 {max(e_all.values()):.2e}; "
            f"timed run, smallest Trotter systems max {max(e_full.values()):.2e}")
    return res


def twodir_leg(args, torch):
    """2DIR with the two-exciton manifold: N=32 ring, Full type, AR(1)
    energies (1650, sd 12 cm^-1), static nearest-neighbour couplings
    N(-6, 2), anharmonicity 16 cm^-1, t1 = t3 = 64 steps, t2 = 0, dt 2 fs,
    400 samples every 10 frames, Trotter 1, all 21 polarisation items."""
    import numpy as np

    from nise_2017_amd import abi, synth
    N, T, sr, ns = 32, 64, 10, 400
    nfr = (ns - 1) * sr + 2 * T + 2
    H, mu = synth.ring_2dir(N=N, nframes=nfr, seed=20261202 + 2)
    lo, hi = synth.shifted_center(H, N)
    mk = lambda end: abi.make_non(singles=N, tmax1=T, tmax2=0, tmax3=T, sample=sr, begin=0, end=end,  # noqa: E731
                                  ts=1, propagation=1, deltat=2.0, min1=lo, max1=hi, technique="2DIR",
                                  anharmonicity=16.0)
    Hd, mud = torch.from_numpy(H).cuda(), torch.from_numpy(mu).cuda()
    abi.twodir_batched(mk(ns), Hd, mud)  # warm: the run's device buffers are kept for the next call
    torch.cuda.synchronize()
    with no_gc():
        t0 = time.perf_counter()
        R = abi.twodir_batched(mk(ns), Hd, mud)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    items = ns * 21
    # the reference's propagation work per item (src/calc_2DIR.c:184-506, t1 = t3 = T): singles vector-steps
    # sum_t1 t1 (GB seeds) + T (GB t3) + (T + 1) T (SE) + (T + 1) T (EA right-hand vectors), and doubles
    # vector-steps (T + 1) T; one site-step per amplitude and frame (10,400 x 32 + 4,160 x 528 at N = 32, T = 64)
    singles = T * (T - 1) // 2 + T + 2 * (T + 1) * T
    ref_site_steps = items * (singles * N + (T + 1) * T * (N * (N + 1) // 2))
    res = {"workload": f"N={N} ring, t1=t3={T}, {ns} samples x 21 items, anharmonicity 16",
           "wall_ms": round(wall * 1e3, 3), "items_per_s": items / wall,
           "reference_equivalent_site_steps_per_s": ref_site_steps / wall,
           "note": "reference_equivalent_site_steps_per_s counts the reference's per-item propagations (each "
                   "item repeats vectors another item shares), not work this run does"}
    log(f"2DIR leg: {wall * 1e3:.1f} ms for {items} work items")
    # 2DIR-Raman on the same system with Raman tensors xx..zz of 0.6 + N(0, 0.3) per site and frame
    rng = np.random.default_rng(20261202 + 3)
    al = np.zeros((nfr, 1 + 6 * N), np.float32)
    al[:, 1:] = (0.6 + 0.3 * rng.normal(size=(nfr, 6 * N))).astype(np.float32)
    ald = torch.from_numpy(al).cuda()
    mkr = lambda end: abi.make_non(singles=N, tmax1=T, tmax2=0, tmax3=T, sample=sr, begin=0, end=end,  # noqa: E731
                                   ts=1, propagation=1, deltat=2.0, min1=lo, max1=hi, technique="2DIRraman",
                                   anharmonicity=16.0)
    abi.twodir_raman_batched(mkr(ns), Hd, mud, ald)
    torch.cuda.synchronize()
    with no_gc():
        t0 = time.perf_counter()
        abi.twodir_raman_batched(mkr(ns), Hd, mud, ald)
        torch.cuda.synchronize()
        wr = time.perf_counter() - t0
    res["raman"] = {"workload": f"2DIRraman, the same system and sizes, {ns} samples x 21 items",
                    "wall_ms": round(wr * 1e3, 3), "items_per_s": items / wr}
    log(f"2DIR-Raman: {wr * 1e3:.1f} ms for {items} work items")
    if not args.no_cpu:
        from oracle import oracle
        k = 2
        t0 = time.perf_counter()
        ref = oracle.twodir(mk(k), H, mu)
        dt = time.perf_counter() - t0
        Rk = abi.twodir_batched(mk(k), Hd, mud)
        err = float(np.linalg.norm(Rk - ref) / np.linalg.norm(ref))
        res["parity_rel_l2_vs_oracle"] = err
        res["cpu_baseline"] = {"value": k * 21 / dt, "unit": "items/s", "cores": 1, "kind": "port",
                               "sample": f"{k} samples x 21 items, oracle item loop, {dt:.2f} s"}
        log(f"2DIR parity {err:.1e}, oracle {k * 21 / dt:.2f} items/s")
        t0 = time.perf_counter()
        refr = oracle.twodir_raman(mkr(k), H, mu, al, acc64=True)
        dtr = time.perf_counter() - t0
        Rr = abi.twodir_raman_batched(mkr(k), Hd, mud, ald)
        res["raman"]["parity_rel_l2_vs_oracle"] = float(np.linalg.norm(Rr - refr) / np.linalg.norm(refr))
        res["raman"]["cpu_baseline"] = {"value": k * 21 / dtr, "unit": "items/s", "cores": 1, "kind": "port",
                                        "sample": f"{k} samples x 21 items, oracle item loop, {dtr:.2f} s"}
        log(f"2DIR-Raman parity {res['raman']['parity_rel_l2_vs_oracle']:.1e}, oracle {k * 21 / dtr:.2f} items/s")
    return res


def host_cpu():
    """(nproc, CPU model name) of this host."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return os.cpu_count() or 1, model


def cpu_threads():
    """Worker threads for the CPU baselines: every host core, up to the 16
    hardware threads a one-GPU box allots to its job."""
    return max(1, min(16, os.cpu_count() or 1))


def cpu_baseline_leg(args, N, T, sr, ns, lo, hi, non, diag, row, mus, H, mu, nfr):
    """The reference's own compiled absorption loop (src/absorption.c:92-146
    over the reference's read_He, read_mue and propagate_vec_coupling_S,
    built from its sources into oracle/_ref) on a bounded sample of the
    same workload, the samples split into contiguous BeginPoint/EndPoint
    ranges, one per thread.  Two modes: frames read with fseek/fread per
    step from files in the page cache (the reference's I/O), and the same
    loop over the trajectory held in memory (stdio streams over RAM
    buffers).  The oracle port is timed beside it.  Returns the
    cpu_baseline object and the GPU's relative L2 against the oracle on the
    same samples."""
    import tempfile
    import threading

    import numpy as np

    from nise_2017_amd import abi
    from oracle import oracle
    threads = cpu_threads()
    nproc, model = host_cpu()
    f_need = lambda k: (k - 1) * sr + T  # noqa: E731

    def make(k):
        n = abi.make_non(singles=N, tmax1=T, sample=sr, begin=0, end=k, ts=1, propagation=1,
                         deltat=2.0, min1=lo, max1=hi)
        n.shifte = non.shifte
        return n

    def ranges(k):
        return [(k * i // threads, k * (i + 1) // threads) for i in range(threads)]

    def run_threads(fn, k):
        parts = [(np.zeros(T, np.float32), np.zeros(T, np.float32)) for _ in range(threads)]
        ths = [threading.Thread(target=fn, args=(b, e, parts[i][0], parts[i][1]))
               for i, (b, e) in enumerate(ranges(k)) if e > b]
        t_a = time.perf_counter()
        for th in ths:
            th.start()
        for th in ths:
            th.join()
        dt = time.perf_counter() - t_a
        re = np.sum([p[0] for p in parts], axis=0)
        im = np.sum([p[1] for p in parts], axis=0)
        return dt, re, im

    out = {"unit": "site-steps/s", "cores": threads, "nproc": nproc, "cpu_model": model}
    ref_ok = oracle.have_ref()
    # per-sample cost on one thread, to size the sample
    Hh, muh = host_records(diag, row, mus, f_need(1))
    if ref_ok:
        R = oracle.ref()
        re1, im1 = np.zeros(T, np.float32), np.zeros(T, np.float32)
        t_a = time.perf_counter()
        R.ref_absorption_mem(N, Hh.ctypes.data, Hh.nbytes, muh.ctypes.data, muh.nbytes, lo, hi, 2.0, 0.0, 1, T,
                             sr, 0, 1, Hh.shape[0], re1, im1)
        per_sample = time.perf_counter() - t_a
    else:
        t_a = time.perf_counter()
        oracle.absorption(make(1), Hh, muh, nthreads=1)
        per_sample = time.perf_counter() - t_a
    # about cpu_seconds of work on every thread, at most 1200 samples (a
    # 1.6 GB trajectory file at N = 256)
    k = int(max(threads, min(ns, 1200, round(args.cpu_seconds * threads / per_sample))))
    Hh, muh = host_records(diag, row, mus, f_need(k))
    site_steps = k * 3 * T * N
    sample_txt = (f"{k} of {ns} samples x 3 pols x {T} steps at N={N}, contiguous sample ranges on "
                  f"{threads} threads")
    if ref_ok:
        shm = "/dev/shm" if os.path.isdir("/dev/shm") else None
        with tempfile.TemporaryDirectory(dir=shm) as d:
            hf, df = os.path.join(d, "H.bin"), os.path.join(d, "mu.bin")
            Hh.tofile(hf)
            muh.tofile(df)

            def files(b, e, r, i):
                R.ref_absorption_files(N, hf.encode(), df.encode(), b"", b"", b"Full", lo, hi, 2.0, 0.0, 1, T,
                                       sr, b, e, Hh.shape[0], r, i)
            run_threads(files, threads)  # page cache warm
            dt_f, rf, imf = run_threads(files, k)

            def mem(b, e, r, i):
                R.ref_absorption_mem(N, Hh.ctypes.data, Hh.nbytes, muh.ctypes.data, muh.nbytes, lo, hi, 2.0, 0.0,
                                     1, T, sr, b, e, Hh.shape[0], r, i)
            dt_m, rm, imm = run_threads(mem, k)
        out.update({"value": site_steps / dt_f, "kind": "reference",
                    "sample": sample_txt + f", the reference's loop compiled from its sources, frames read "
                                          f"per step from files in the page cache, {dt_f:.1f} s",
                    "modes": {"file_read": {"value": site_steps / dt_f, "seconds": round(dt_f, 2)},
                              "preloaded": {"value": site_steps / dt_m, "seconds": round(dt_m, 2)}},
                    "single_thread_value": 3 * T * N / per_sample})
        log(f"cpu baseline (reference loop, {threads} threads): files {site_steps / dt_f:.3e}, "
            f"memory {site_steps / dt_m:.3e} site-steps/s")
    # the oracle port on the same samples (and the GPU's parity against it)
    t_a = time.perf_counter()
    o_re, o_im = oracle.absorption(make(k), Hh, muh, nthreads=threads)
    dt_o = time.perf_counter() - t_a
    port = {"value": site_steps / dt_o, "seconds": round(dt_o, 2), "kind": "port", "cores": threads}
    if ref_ok:
        out["port"] = port
        out["reference_vs_port_rel_l2"] = float(np.linalg.norm((rf - o_re) + 1j * (imf - o_im)) /
                                                np.linalg.norm(o_re + 1j * o_im))
    else:
        out.update({"value": port["value"], "kind": "port", "sample": sample_txt + ", trajectory preloaded"})
    pc = abi.AbsPlan(make(k), H, H.shape[1], mu, mu.shape[1], nfr, where=abi.NISE_DEVICE)
    g_re, g_im = pc.execute()
    pc.close()
    err = float(np.linalg.norm((g_re - o_re) + 1j * (g_im - o_im)) / np.linalg.norm(o_re + 1j * o_im))
    return out, err


def chain16_leg(args, torch):
    """The latency-bound case: N=16 chain, Full type, AR(1) energies (1650,
    sd 10, 100 fs) and nearest-neighbour couplings N(-5, 2) + AR(1) sd 1,
    4200 frames at dt 2 fs, tmax 128, 200 samples every 20 frames, 3
    polarisations, Trotter 1: the whole run as one captured graph, and the
    plain per-sample kernel on the same run."""
    import numpy as np

    from nise_2017_amd import abi, synth
    N, T, sr, ns = 16, 128, 20, 200
    H, mu = synth.chain_absorption(N=N, nframes=4200, seed=20261202)
    lo, hi = synth.shifted_center(H, N)
    non = abi.make_non(singles=N, tmax1=T, sample=sr, begin=0, end=ns, ts=1, propagation=1,
                       deltat=2.0, min1=lo, max1=hi, fft=2048, lifetime=1300.0)
    non.shifte = np.float32((np.float32(hi) + np.float32(lo)) / np.float32(2))
    Hd, mud = torch.from_numpy(H).cuda(), torch.from_numpy(mu).cuda()
    plan = abi.AbsPlan(non, Hd, H.shape[1], mud, mu.shape[1], H.shape[0], where=abi.NISE_DEVICE)
    for _ in range(3):
        plan.execute(flags=2, fetch=False)
    torch.cuda.synchronize()
    reps = 50
    with no_gc():
        t0 = time.perf_counter()
        for _ in range(reps):
            plan.execute(flags=2, fetch=False)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / reps
    plan.execute(flags=1, fetch=False)
    t = plan.timing()
    info = plan.info()
    plan.close()
    pp = abi.AbsPlan(non, Hd, H.shape[1], mud, mu.shape[1], H.shape[0], where=abi.NISE_DEVICE, kernel=1)
    pp.execute(flags=1, fetch=False)
    pms = pp.timing()["conveyor_ms"]
    pp.close()
    site_steps = ns * 3 * T * N
    res = {"workload": "absorption N=16 chain, Full FP32 frames, 4200 frames, tmax 128, 200 samples every 20 "
                       "frames, 3 pols, Trotter 1",
           "ms_per_run": round(ms, 4), "site_steps_per_s": site_steps / (ms * 1e-3),
           "stages_ms": {k: round(v, 4) for k, v in t.items() if k != "adjoint_ms"},
           "rows": info["rows"], "plain_kernel_ms": round(pms, 3),
           "batched_over_plain": round(pms / ms, 2)}
    log(f"N=16 chain: {ms:.3f} ms per run ({res['site_steps_per_s']:.3e} site-steps/s), plain {pms:.2f} ms")
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--samples", type=int, default=4000)
    ap.add_argument("--frames", type=int, default=50000)
    ap.add_argument("--lattice", type=int, default=16)
    ap.add_argument("--tmax", type=int, default=256)
    ap.add_argument("--samplerate", type=int, default=10)
    ap.add_argument("--plain-samples", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--shape", default="", help="vp,spw,block conveyor shape override")
    ap.add_argument("--rows", type=int, default=0, help="row conveyor with this many rows")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the N=128 adjoint leg")
    ap.add_argument("--no-multi", action="store_true", help="skip the uneven-size batch leg")
    ap.add_argument("--no-2dir", action="store_true", help="skip the 2DIR leg")
    ap.add_argument("--no-cfg0", action="store_true", help="skip the N=16 chain leg")
    ap.add_argument("--multi-systems", type=int, default=96)
    ap.add_argument("--adjoint-every", default="16,32,64", help="checkpoint intervals K")
    ap.add_argument("--fd-params", type=int, default=3, help="offsets checked by finite differences")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first copy 1 GiB on the device (a kernel of known traffic for scaling "
                         "FETCH_SIZE / WRITE_SIZE counter passes)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(one GPU's share of the sample range is the unit that would shard)", file=sys.stderr)
        sys.exit(2)

    import ctypes as C

    import numpy as np
    import torch

    from nise_2017_amd import abi

    torch.cuda.set_device(0)
    if args.copy_probe:  # 1 GiB read + 1 GiB written by one copy kernel
        src = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    L, nfr, T, sr, ns = args.lattice, args.frames, args.tmax, args.samplerate, args.samples
    N = L * L
    need = (ns - 1) * sr + T
    if need > nfr:
        raise SystemExit(f"{ns} samples need {need} frames, have {nfr}")
    log(f"generating {nfr} frames, N={N} on the device")
    H, mu, diag, coup, mus, row = build_lattice_on_device(L, nfr, 20261202 + 1, torch)
    mean = float(diag[:need].mean())
    lo, hi = mean - 100.0, mean + 100.0
    non = abi.make_non(singles=N, tmax1=T, sample=sr, begin=0, end=ns, ts=1, propagation=1,
                       deltat=2.0, min1=lo, max1=hi, couplingcut=0.0, fft=2048)
    non.shifte = np.float32((np.float32(hi) + np.float32(lo)) / np.float32(2))
    torch.cuda.synchronize()

    log("building the plan (coupling structure, rotation schedule)")
    plan = abi.AbsPlan(non, H, H.shape[1], mu, mu.shape[1], nfr, where=abi.NISE_DEVICE, kernel=0)
    tune = None
    if args.shape:
        vp, spw, blk = (int(x) for x in args.shape.split(","))
        plan.set_shape(vp, spw, blk)
    elif args.rows:
        plan.set_rows(args.rows)
    else:
        tune = plan.autotune()
        log("autotune (code, conveyor ms): " + ", ".join(f"({c}, {t:.2f})" for c, t in tune))
    info = plan.info()
    log(f"plan: {info}")

    # per-kernel timing with HIP events on the plan's stream
    plan.execute(flags=1, fetch=False)
    conv_ms, tab_ms = [], []
    for _ in range(max(1, min(3, args.steps))):
        plan.execute(flags=1, fetch=False)
        t = plan.timing()
        conv_ms.append(t["conveyor_ms"])
        tab_ms.append(t["tables_ms"])
    log(f"event timing: tables {np.mean(tab_ms):.3f} ms, conveyor {np.mean(conv_ms):.3f} ms")

    for _ in range(args.warmup):
        plan.execute(flags=2, fetch=False)
    torch.cuda.synchronize()
    with no_gc():
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan.execute(flags=2, fetch=False)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    re_g, im_g = plan.execute(flags=0)
    site_steps = ns * 3 * T * N
    value = site_steps / (ms_per_step * 1e-3)
    log(f"{ms_per_step:.3f} ms per run, {value:.3e} site-steps/s")

    B = algorithmic_bytes_per_site_step(N, T, sr)
    conv = float(np.mean(conv_ms))
    achieved = site_steps * B / (conv * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "rows_kernel" if info["rows"] else "conveyor_kernel", "kernel_ms": round(conv, 4),
                "bytes_per_site_step": round(B, 3)}
    # HBM bytes per conveyor launch from rocprofv3 FETCH_SIZE / WRITE_SIZE
    # passes over the same workload (scaled by a copy kernel of known size)
    tf = os.path.join(ROOT, "profiles", f"{roofline['kernel']}_lattice256_traffic.json")
    if os.path.exists(tf):
        with open(tf) as fh:
            roofline["traffic"] = json.load(fh).get("bytes_per_launch")
        roofline["traffic_source"] = (f"profiles/{os.path.basename(tf)}: rocprofv3 FETCH_SIZE and WRITE_SIZE "
                                      "passes over this workload, scaled by a copy kernel (not measured in "
                                      "this run)")

    extra = {}
    if not args.no_plain and args.plain_samples > 0:
        nps = args.plain_samples
        nonp = abi.make_non(singles=N, tmax1=T, sample=sr, begin=0, end=nps, ts=1, propagation=1,
                            deltat=2.0, min1=lo, max1=hi)
        nonp.shifte = non.shifte
        pp = abi.AbsPlan(nonp, H, H.shape[1], mu, mu.shape[1], nfr, where=abi.NISE_DEVICE,
                         kernel=1)
        pp.execute(flags=1, fetch=False)
        pms = pp.timing()["conveyor_ms"]
        pp.close()
        plain_rate = nps * 3 * T * N / (pms * 1e-3)
        # frame-major on the same subset, for a like-for-like ratio
        pf = abi.AbsPlan(nonp, H, H.shape[1], mu, mu.shape[1], nfr, where=abi.NISE_DEVICE,
                         kernel=0)
        pf.execute(flags=1, fetch=False)
        pf.execute(flags=1, fetch=False)
        fms = sum(pf.timing()[k] for k in ("tables_ms", "conveyor_ms", "reduce_ms"))
        pf.close()
        extra["plain_kernel"] = {"value": plain_rate, "samples": nps, "ms": round(pms, 3),
                                 "batched_same_subset_ms": round(fms, 3),
                                 "batched_over_plain_full_run": round(value / plain_rate, 2)}
        log(f"plain kernel: {plain_rate:.3e} site-steps/s on {nps} samples")

    cpu = None
    if not args.no_cpu:
        cpu, extra["parity_rel_l2_vs_fp32_oracle"] = cpu_baseline_leg(args, N, T, sr, ns, lo, hi, non, diag, row,
                                                                      mus, H, mu, nfr)
    if not args.no_cfg0:
        extra["chain16"] = chain16_leg(args, torch)

    if not args.no_adjoint:
        extra["adjoint"] = adjoint_leg(args, torch)
    if not args.no_multi:
        extra["multi"] = multi_leg(args, torch)
    if not args.no_2dir:
        extra["twodir"] = twodir_leg(args, torch)

    line = {
        "metric": "site-steps/s propagated (N=256 absorption) vs plain GPU; adjoint/forward cost",
        "value": value, "unit": "site-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"absorption N={N} {L}x{L} dipole lattice, Full FP32 frames, "
                               f"{nfr} frames, tmax {T}, {ns} samples every {sr} frames, 3 pols, "
                               f"Trotter 1", "N": N, "frames": nfr, "tmax": T,
                   "samples": ns, "samplerate": sr,
                   "conveyor": ({"kernel": "rows_kernel", "rows": info["rows"]} if info["rows"] else
                                {"kernel": "conveyor_kernel", "vp": info["vp"], "spw": info["spw"],
                                 "block": info["block"]}),
                   "rotations": info["rotations"], "layers": info["nlayers"],
                   "autotune": [{"code": c, "ms": round(t, 3)} for c, t in tune] if tune else None},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    line.update(extra)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
