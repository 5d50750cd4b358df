"""Gradients of the IR lineshape with respect to the map coefficients.

The forward pass is the calcIR IR channel (field, maps and Hamiltonian of
calcIR.cu:1020-1404, propagation and mu.F.mu correlation of
calcIR.cu:577-648, decay and C2R of calcIR.cu:777-827) run by libcalcir's
cir_forward_ls; the reverse pass is cir_backward_ls (adjoint propagation with
the Daleckii-Krein derivative of exp(iH dt/hbar), contracted with dH/dcoef
and dmu/dcoef on the device).  The reference computes no derivatives.

    sess = LineshapeSession(p, xyz, box, times)
    ft = sess.forward(coefs)                 # C2R of the IR TCF (nfft values)
    g = sess.backward(dL_dft)                # dL/dcoef, 13 values
    ls = IRLineshape.apply(coefs_tensor, sess)   # torch.autograd form
"""

import ctypes
import functools

import numpy as np

from . import NCOEF, check, check_trajectory, lib

__all__ = ["LineshapeSession", "IRLineshape", "lineshape_map", "lineshape_rows"]


def lineshape_map(p):
    """Rows of <outf>_irls.dat as (index into the C2R output, frequency,
    scale): value = ftcf[index] * scale (calcIR.cu:893-919, float arithmetic,
    the second loop bound with the reference's operator precedence)."""
    nt, nz = p.ntcfpoints, p.nzeros
    n = nt + nz
    factor = np.float32(2 * 3.14159265359 * 5.308837367 / (np.float32(p.dt) * np.float32(n)))
    idx, fr = [], []
    for i in range(n // 2, n):
        f = np.float32(-(i - nt - nz)) * factor + np.float32(p.avef)
        if f <= p.omegaStop:
            idx.append(i)
            fr.append(f)
    for i in range(0, nt + nz // 2):
        f = np.float32(-i) * factor + np.float32(p.avef)
        if f >= p.omegaStart:
            idx.append(i)
            fr.append(f)
    scale = 1.0 / (float(factor) * n)
    return np.array(idx, np.int64), np.array(fr, np.float64), scale


def lineshape_rows(p, ftcf):
    """(frequency, value) of the IR lineshape rows for a C2R output."""
    idx, fr, scale = lineshape_map(p)
    return fr, np.asarray(ftcf)[idx] * scale


class LineshapeSession:
    """One context holding a trajectory for repeated forward/backward passes
    at different map coefficients."""

    def __init__(self, p, xyz, box, times, stream=None):
        self.p = p
        self.xyz, self.box, self.times = check_trajectory(p, xyz, box, times)
        self.nfft = p.ntcfpoints + p.nzeros
        # forward passes run so far and the coefficients of the last one:
        # cir_backward_ls differentiates the context's last forward pass
        self.generation = 0
        self.forward_coefs = None
        self._ctx = ctypes.c_void_p()
        check(lib().cir_create(ctypes.byref(p), self.xyz.shape[1], stream, ctypes.byref(self._ctx)))

    def close(self):
        if self._ctx:
            lib().cir_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def coefs(self):
        c = np.zeros(NCOEF)
        check(lib().cir_get_coefs(self._ctx, c.ctypes.data))
        return c

    def forward(self, coefs=None):
        """C2R output of the IR correlation function at these coefficients
        (None keeps the context's current set)."""
        if coefs is not None:
            cf = np.ascontiguousarray(coefs, np.float64)
            if cf.shape != (NCOEF,):
                raise ValueError("coefs must hold %d values" % NCOEF)
            check(lib().cir_set_coefs(self._ctx, cf.ctypes.data))
        ft = np.zeros(self.nfft)
        self.forward_coefs = None
        check(lib().cir_forward_ls(self._ctx, self.xyz.shape[0], self.xyz.ctypes.data, self.box.ctypes.data,
                                   self.times.ctypes.data, ft.ctypes.data))
        self.generation += 1
        self.forward_coefs = self.coefs()
        return ft

    def backward(self, dL_dft):
        """dL/dcoef at the coefficients of the last forward pass."""
        g = np.ascontiguousarray(dL_dft, np.float64)
        if g.shape != (self.nfft,):
            raise ValueError("dL_dft must hold %d values" % self.nfft)
        out = np.zeros(NCOEF)
        check(lib().cir_backward_ls(self._ctx, g.ctypes.data, out.ctypes.data))
        return out


@functools.lru_cache(maxsize=None)
def _function():
    import torch

    class IRLineshape(torch.autograd.Function):
        """IR lineshape rows (as written to <outf>_irls.dat) as a function of
        the 13 map coefficients (float64 tensor), differentiable through
        cir_backward_ls."""

        @staticmethod
        def forward(ctx, coefs, session):
            cf = coefs.detach().cpu().numpy().astype(np.float64)
            ft = session.forward(cf)
            idx, _, scale = lineshape_map(session.p)
            ctx.session = session
            ctx.coefs = cf
            ctx.generation = session.generation
            ctx.device = coefs.device
            ctx.idx, ctx.scale = idx, scale
            return torch.as_tensor(ft[idx] * scale, dtype=torch.float64, device=coefs.device)

        @staticmethod
        def backward(ctx, grad_out):
            sess = ctx.session
            # the session differentiates its last forward pass: if another
            # forward ran since this one, run this one's again first
            if sess.generation != ctx.generation or sess.forward_coefs is None or \
                    not np.array_equal(sess.forward_coefs, ctx.coefs):
                sess.forward(ctx.coefs)
                ctx.generation = sess.generation
            g = np.zeros(sess.nfft)
            np.add.at(g, ctx.idx, grad_out.detach().cpu().numpy().astype(np.float64) * ctx.scale)
            grad = torch.as_tensor(sess.backward(g), dtype=torch.float64)
            return grad.to(ctx.device), None

    return IRLineshape


class _Lazy:
    def __getattr__(self, name):
        return getattr(_function(), name)


IRLineshape = _Lazy()
