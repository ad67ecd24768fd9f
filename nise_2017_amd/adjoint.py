"""Gradients of the linear absorption spectrum through the GPU adjoint.

The spectrum do_1DFFT writes (src/1DFFT.c:144-228) is linear in the
accumulated response S(t) = re_S + i im_S:

    spec_r[k] = 2 dt c_v sum_n w_n (re_S[n] cos(2 pi n i_k / F) - im_S[n] sin(2 pi n i_k / F))

with w_n = damping(n) / samples, w_0 halved, F the FFT length and i_k the
FFT bins inside (min1, max1) in file order.  spectrum_matrix returns that
map; the loss and its gradient with respect to S are ordinary torch (or
numpy) algebra, and AbsorptionResponse carries dL/dS back to the per-site
energy offsets and the coupling scale with nise_abs_plan_adjoint.
"""
import numpy as np

from . import abi

C_V = np.float32(2.99792458e-5)


def spectrum_matrix(non, samples):
    """(freq, A_re, A_im): spec_r = A_re @ re_S + A_im @ im_S (float64)."""
    T = non.tmax1
    fft = max(non.fft, T)
    dt = np.float32(non.deltat)
    n = np.arange(T, dtype=np.float64)
    w = np.full(T, 1.0 / samples)
    if non.lifetime > 0:
        w *= np.exp(-n * float(dt) / (2 * float(np.float32(non.lifetime))))
    if non.homogen > 0:
        w *= np.exp(-n * float(dt) / (2 * float(np.float32(non.homogen))))
    if non.inhomogen > 0:
        ih = float(np.float32(non.inhomogen))
        w *= np.exp(-n * float(dt) * n * float(dt) / (2 * ih * ih))
    if non.window == 1:
        w *= np.cos(n * 2 * np.pi / (4 * non.tmax - 4)) ** 2
    w[0] *= 0.5
    shift1 = np.float32(non.shifte)
    bins, freq = [], []
    order = list(range(fft // 2, fft)) + list(range(0, fft // 2))
    for i in order:
        num = np.float32(fft - i) if i >= fft // 2 else np.float32(-i)
        wv = -(num / dt / C_V / np.float32(fft) - shift1)
        if non.min1 < wv < non.max1:
            bins.append(i)
            freq.append(float(wv))
    bins = np.array(bins, dtype=np.int64)
    ph = 2 * np.pi * ((np.outer(bins, np.arange(T)) % fft) / fft)
    scale = 2.0 * float(dt) * float(C_V)
    A_re = scale * np.cos(ph) * w[None, :]
    A_im = -scale * np.sin(ph) * w[None, :]
    return np.array(freq), A_re, A_im


class AdjointAbsorption:
    """A batched absorption plan with gradient parameters and checkpoints.

    forward(delta, cscale) -> (re_S, im_S) float64 arrays
    backward(g_re, g_im)   -> (dL/d delta, dL/d cscale)
    """

    def __init__(self, non, H, mu, nframes=None, coupling_rec=None, every=16, where=None,
                 shape=None, rows=None):
        if where is None:
            where = abi.NISE_DEVICE if hasattr(H, "data_ptr") and H.is_cuda else abi.NISE_HOST
        nfr = H.shape[0] if nframes is None else nframes
        self.plan = abi.AbsPlan(non, H, H.shape[1], mu, mu.shape[1], nfr, coupling_rec, where=where)
        if shape is not None:
            self.plan.set_shape(*shape)
        elif rows is not None:
            self.plan.set_rows(rows)
        self.plan.enable_adjoint(every)
        self.N, self.T = non.singles, non.tmax1

    def forward(self, delta=None, cscale=1.0, timed=False):
        self.plan.set_params(delta, cscale)
        re, im = self.plan.execute(flags=1 if timed else 0)
        return re.astype(np.float64), im.astype(np.float64)

    def backward(self, g_re, g_im):
        return self.plan.adjoint(g_re, g_im)

    def timing(self):
        return self.plan.timing()

    def close(self):
        self.plan.close()


def _autograd():
    import torch

    class AbsorptionResponse(torch.autograd.Function):
        """S = (re_S, im_S) as a function of (delta, cscale); backward runs
        the GPU adjoint of the plan's last forward."""

        @staticmethod
        def forward(ctx, delta, cscale, model):
            c = float(cscale)
            re, im = model.forward(delta.detach().cpu().numpy(), c)
            ctx.model = model
            # gradients go back on the device and in the dtype of each input;
            # a plain number as the scale takes no gradient
            ctx.d_meta = (delta.device, delta.dtype)
            ctx.c_meta = (cscale.device, cscale.dtype) if torch.is_tensor(cscale) else None
            return torch.from_numpy(np.stack([re, im])).to(device=delta.device, dtype=delta.dtype)

        @staticmethod
        def backward(ctx, gS):
            g = gS.detach().to("cpu", torch.float64).numpy()
            gd, gc = ctx.model.backward(g[0], g[1])
            dev, dt = ctx.d_meta
            gdt = torch.from_numpy(np.asarray(gd, np.float64)).to(device=dev, dtype=dt)
            gct = None
            if ctx.c_meta is not None:
                gct = torch.tensor(float(gc), dtype=torch.float64).to(device=ctx.c_meta[0], dtype=ctx.c_meta[1])
            return gdt, gct, None

    return AbsorptionResponse


def absorption_response(delta, cscale, model):
    """torch: S(delta, cscale) of shape (2, tmax1) with a GPU adjoint."""
    return _autograd().apply(delta, cscale, model)
