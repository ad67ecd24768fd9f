"""GPU adjoint of the absorption spectrum (nise_abs_plan_adjoint) against
central finite differences of the FP64 oracle.

Loss: L = |spec_r - target|^2 over the do_1DFFT grid, spec_r = A_re re_S +
A_im im_S (spectrum_matrix), target from other offsets and scale.  The
gradient with respect to the per-site offsets (h = 0.01 cm^-1) and the
coupling scale (h = 1e-4) must agree to 1e-3 relative; the forward with
delta = 0, c = 1 must equal the plain batched absorption bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

from nise_2017_amd import abi, adjoint, synth

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[adjoint] {msg}", flush=True)


def make_case(N=8, T=48, sr=7, ns=12, seed=61, ts=1, ctype=False):
    if ctype:
        L = int(round(np.sqrt(N)))
        rng, mus, coup = synth.lattice_static(L, seed=seed, cut=0.1)
        nfr = (ns - 1) * sr + T + 3
        Hf = synth.full_records(synth.lattice_diag(rng, nfr, N), coup)
        a, b = np.triu_indices(N)
        H = np.ascontiguousarray(np.concatenate([Hf[:, :1], Hf[:, 1:][:, a == b]], axis=1))
        Cr = np.ascontiguousarray(Hf[0])
        mu = synth.dipole_records(mus, 1)
        lo, hi = synth.shifted_center(Hf, N)
        extra = dict(hamiltonian="Coupling")
    else:
        H, mu = synth.chain_absorption(N=N, nframes=(ns - 1) * sr + T + 3, seed=seed)
        Cr = None
        lo, hi = synth.shifted_center(H, N)
        extra = {}
    non = abi.make_non(singles=N, tmax1=T, sample=sr, begin=0, end=ns, ts=ts, propagation=1,
                       deltat=2.0, min1=lo, max1=hi, fft=256, lifetime=300.0, **extra)
    non.shifte = np.float32((np.float32(hi) + np.float32(lo)) / np.float32(2))
    return non, H, mu, Cr


def loss_grad(S, A_re, A_im, target):
    re, im = S
    r = A_re @ re + A_im @ im - target
    return float(r @ r), 2 * A_re.T @ r, 2 * A_im.T @ r


def fd_grad(orc, non, H, mu, Cr, d0, c0, A_re, A_im, target, h=0.01, hc=1e-4):
    def L(d, c):
        return loss_grad(orc.absorption_param_f64(non, H, mu, d, c, Cr, nthreads=8), A_re, A_im,
                         target)[0]
    g = np.zeros(len(d0))
    for a in range(len(d0)):
        e = np.zeros(len(d0))
        e[a] = h
        g[a] = (L(d0 + e, c0) - L(d0 - e, c0)) / (2 * h)
    gc = (L(d0, c0 + hc) - L(d0, c0 - hc)) / (2 * hc)
    return g, gc


def run_case(orc, every, shape=None, tol=1e-4, rows=None, **kw):
    """dL/dS from the FP64 oracle's S isolates the adjoint from the FP32
    forward error: the residual spec - target is a few 1e-3 of spec here, so
    the forward's ~1e-6 relative error would otherwise move dL/dS by ~1e-3
    (test_end_to_end covers the GPU's own dL/dS)."""
    non, H, mu, Cr = make_case(**kw)
    rng = np.random.default_rng(kw.get("seed", 61) + 1)
    N = non.singles
    d0 = rng.normal(0, 2, N).astype(np.float32).astype(np.float64)
    c0 = float(np.float32(1.05))
    freq, A_re, A_im = adjoint.spectrum_matrix(non, non.end)
    St = orc.absorption_param_f64(non, H, mu, rng.normal(0, 5, N), 1.1, Cr)
    target = A_re @ St[0] + A_im @ St[1]
    model = adjoint.AdjointAbsorption(non, H, mu, coupling_rec=Cr, every=every, shape=shape,
                                      rows=rows)
    S = model.forward(d0, c0)
    So = orc.absorption_param_f64(non, H, mu, d0, c0, Cr)
    err_fwd = np.linalg.norm((S[0] - So[0]) + 1j * (S[1] - So[1])) / np.linalg.norm(So[0] + 1j * So[1])
    assert err_fwd < 1e-4
    _, g_re, g_im = loss_grad(So, A_re, A_im, target)
    gd, gc = model.backward(g_re, g_im)
    fd, fdc = fd_grad(orc, non, H, mu, Cr, d0, c0, A_re, A_im, target)
    model.close()
    e_d = np.linalg.norm(gd - fd) / np.linalg.norm(fd)
    e_c = abs(gc - fdc) / abs(fdc)
    say(f"every={every} shape={shape} kw={kw}: fwd {err_fwd:.1e}, d-grad {e_d:.1e}, c-grad {e_c:.1e}")
    assert e_d < tol and e_c < tol
    return gd, gc


def test_end_to_end(orc):
    """The GPU's own forward gives dL/dS; target far from the start."""
    non, H, mu, Cr = make_case(N=8, T=48, seed=67)
    N = non.singles
    rng = np.random.default_rng(68)
    freq, A_re, A_im = adjoint.spectrum_matrix(non, non.end)
    St = orc.absorption_param_f64(non, H, mu, rng.normal(0, 30, N), 1.6, Cr)
    target = A_re @ St[0] + A_im @ St[1]
    model = adjoint.AdjointAbsorption(non, H, mu, every=16)
    S = model.forward(None, 1.0)
    L0, g_re, g_im = loss_grad(S, A_re, A_im, target)
    gd, gc = model.backward(g_re, g_im)
    model.close()
    fd, fdc = fd_grad(orc, non, H, mu, Cr, np.zeros(N), 1.0, A_re, A_im, target)
    g = np.append(gd, gc)
    f = np.append(fd, fdc)
    err = np.linalg.norm(g - f) / np.linalg.norm(f)
    say(f"end to end: gradient rel err {err:.1e}, c-grad {abs(gc - fdc) / abs(fdc):.1e}")
    assert err < 1e-3


@pytest.mark.parametrize("every", [1, 8, 47, 64])
def test_gradient_vs_fd_chain(orc, every):
    run_case(orc, every)


def test_gradient_trotter2(orc):
    run_case(orc, 8, ts=2, N=6, T=32)


def test_gradient_coupling_type(orc):
    run_case(orc, 16, N=9, T=40, ctype=True, seed=63)


@pytest.mark.parametrize("shape", [(16, 5, 256), (8, 2, 64), (16, 5, -4), (8, 2, -2)])
def test_checkpoints_in_every_conveyor_variant(orc, shape):
    run_case(orc, 4, shape=shape, N=8, T=40, seed=64)


@pytest.mark.parametrize("rows", [1, 4, 16])
def test_row_adjoint_variants(orc, rows):
    run_case(orc, 8, rows=rows, N=9, T=40, seed=69)
    run_case(orc, 5, rows=rows, N=9, T=24, seed=70, ctype=True)


def test_forward_without_parameters_is_bit_exact():
    say("bit exactness at delta = 0, c = 1")
    non, H, mu, Cr = make_case(N=10, T=40, seed=65)
    L = abi.lib()
    re = np.zeros(40, np.float32)
    im = np.zeros(40, np.float32)
    abi.check(L.nise_absorption_batched(C.byref(non), abi.ptr(H), H.shape[1], abi.ptr(mu), mu.shape[1],
                                        H.shape[0], None, abi.NISE_HOST, 0, re, im))
    model = adjoint.AdjointAbsorption(non, H, mu, every=8)
    r2, i2 = model.forward(None, 1.0)
    r3, i3 = model.forward(np.zeros(10, np.float32), 1.0)
    model.close()
    assert np.array_equal(r2, re) and np.array_equal(i2, im)
    assert np.array_equal(r3, re) and np.array_equal(i3, im)


def test_autograd_harness(orc):
    import torch
    say("torch.autograd harness")
    non, H, mu, Cr = make_case(N=6, T=32, seed=66)
    freq, A_re, A_im = adjoint.spectrum_matrix(non, non.end)
    target = torch.from_numpy(A_re @ orc.absorption_param_f64(non, H, mu, np.full(6, 1.5), 1.1)[0])
    model = adjoint.AdjointAbsorption(non, H, mu, every=8)
    d = torch.zeros(6, dtype=torch.float64, requires_grad=True)
    c = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    S = adjoint.absorption_response(d, c, model)
    spec = torch.from_numpy(A_re) @ S[0] + torch.from_numpy(A_im) @ S[1]
    loss = ((spec - target) ** 2).sum()
    loss.backward()
    _, g_re, g_im = loss_grad((S[0].detach().numpy(), S[1].detach().numpy()), A_re, A_im,
                              target.numpy())
    gd, gc = model.backward(g_re, g_im)
    model.close()
    assert np.allclose(d.grad.numpy(), gd, rtol=1e-6, atol=0)
    assert abs(c.grad.item() - gc) <= 1e-6 * abs(gc)


def test_autograd_harness_cuda_float32(orc):
    """Parameters on the GPU in float32, and a plain-number scale: the
    gradient comes back on the parameters' device and dtype; the number
    takes none."""
    import torch
    say("torch.autograd harness, cuda float32 parameters")
    non, H, mu, Cr = make_case(N=6, T=32, seed=67)
    freq, A_re, A_im = adjoint.spectrum_matrix(non, non.end)
    target = torch.from_numpy(A_re @ orc.absorption_param_f64(non, H, mu, np.full(6, 1.5), 1.1)[0])
    model = adjoint.AdjointAbsorption(non, H, mu, every=8)
    d = torch.zeros(6, dtype=torch.float32, device="cuda", requires_grad=True)
    S = adjoint.absorption_response(d, 1.0, model)
    assert S.device.type == "cuda" and S.dtype == torch.float32
    spec = torch.from_numpy(A_re).cuda() @ S[0].double() + torch.from_numpy(A_im).cuda() @ S[1].double()
    loss = ((spec - target.cuda()) ** 2).sum()
    loss.backward()
    assert d.grad.device.type == "cuda" and d.grad.dtype == torch.float32
    _, g_re, g_im = loss_grad((S[0].detach().double().cpu().numpy(), S[1].detach().double().cpu().numpy()),
                              A_re, A_im, target.numpy())
    gd, _ = model.backward(g_re, g_im)
    model.close()
    assert np.allclose(d.grad.cpu().numpy(), gd.astype(np.float32), rtol=1e-5, atol=1e-12)
